"""CPU tests of the reference's run files: the parameter deck parser
against the reference's own loadconf (load3mg.c:2-180) on a deck in its
format (golden g13: the values loadconf read from g13_mode.t3c with
g13_file.t3c), and the .prn state file against the reference's saver
(load3mg.c:898-1096): reading its file gives the state it was written from,
and writing that state again gives the same bytes."""
import os

import numpy as np
import pytest

import i3elvis_amd as I
from conftest import GOLDEN, golden


def test_mode_deck_matches_reference_loadconf():
    g = golden("g13_io.npz")
    m = I.read_mode(os.path.join(GOLDEN, "g13_mode.t3c"), os.path.join(GOLDEN, "g13_file.t3c"))
    assert m["fln3"] == int(g["fln3"]) and m["fl1in"] == str(g["fl1in"])
    assert m["fl0num"] == int(g["fl0num"]) and m["fl0out"] == list(g["fl0out"])
    assert np.array_equal(m["fl0otp"], g["fl0otp"]) and np.array_equal(m["fl0cyc"], g["fl0cyc"])
    assert np.array_equal(m["fl0stp"], g["fl0stp"])  # incl. the years -> seconds step
    assert np.array_equal(m["multinnn"], g["multinnn"]) and np.array_equal(m["multittt"], g["multittt"])
    for n, v in zip(g["int_names"], g["int_vals"]):
        if n in m:
            assert m[n] == v, n
    for n, v in zip(g["dbl_names"], g["dbl_vals"]):
        assert m[str(n)] == v, n


def test_mode_deck_errors(tmp_path):
    g = open(os.path.join(GOLDEN, "g13_mode.t3c")).read()
    p = tmp_path / "short.t3c"
    p.write_text(g[: len(g) // 2])
    with pytest.raises(I.I3GError):
        I.read_mode(p)
    with pytest.raises(I.I3GError):
        I.read_mode(os.path.join(GOLDEN, "g13_mode.t3c"), cap=2)  # three output files
    # a run number beyond the table keeps the deck's data file
    f = tmp_path / "file.t3c"
    f.write_text("9 \n")
    assert I.read_mode(os.path.join(GOLDEN, "g13_mode.t3c"), f)["fln3"] == -1


def test_prn_read_matches_reference_saver_and_write_is_byte_identical(tmp_path):
    g = golden("g13_io.npz")
    path = os.path.join(GOLDEN, "g13_state.prn")
    d = I.read_prn(path)
    for k in ("xnumx", "ynumy", "znumz", "mnumx", "mnumy", "mnumz", "pxinit", "pyinit", "pzinit", "rocknum",
              "marknum"):
        assert d[k] == int(g["prn_" + k]), k
    for k, gk in (("xsize", "xsize"), ("ysize", "ysize"), ("zsize", "zsize"), ("pinit", "pinit"),
                  ("gxkoef", "GXKOEF"), ("gykoef", "GYKOEF"), ("gzkoef", "GZKOEF"), ("gridcur", "gridcur"),
                  ("gridtot", "gridtot")):
        assert d[k] == float(g["prn_" + gk]), k
    assert d["timesum_years"] == float(g["prn_timesum"]) / 3.15576e7 and d["n0"] == int(g["prn_n0"])
    assert np.array_equal(d["rock"], g["prn_rock"]) and np.array_equal(d["node"], g["prn_node"])
    for a in ("gx", "gy", "gz"):
        assert np.array_equal(d[a], g["prn_" + a])
    assert np.array_equal(d["bond_at"], g["prn_bond_at"]) and np.array_equal(d["bond_rec"], g["prn_bond_rec"])
    assert np.array_equal(d["bond_v"], g["prn_bond_v"]) and np.array_equal(d["bond_partner"], g["prn_bond_partner"])
    assert np.array_equal(d["mark"], g["prn_mark"]) and np.array_equal(d["markt"], g["prn_markt"])
    out = tmp_path / "again.prn"
    I.write_prn(out, d)
    assert out.read_bytes() == open(path, "rb").read()
    # a truncated file is refused
    bad = tmp_path / "cut.prn"
    bad.write_bytes(open(path, "rb").read()[:-5])
    with pytest.raises(I.I3GError):
        I.read_prn(bad)
