"""GPU parity of the whole marker rheology and movemark's motion, through
the C ABI, against the reference's own movemark and ronurecalc run with the
rock table of its deck (goldens g14, g15: power law, brittle strength,
Karato & Wu, the grain-size model; outgrid 1/2, markmod 0/1, fluid and
immobile markers; retyping, water/air mixing, phase densities; g17: the
partially molten rocks of the melt types).  On these
fixtures every output is equal to the reference's to the bit; the bounds
allow for the device's own exp/pow/asin on other inputs: marker viscosities,
grain roughness and dislocation shares within 2e-6 relative (the grain
growth repeats up to 50 exponential steps), strain to 1e-6 absolute with
the same sign, positions within 2 ulp, types exact; the node fields of
the transfer within 2e-6 relative."""
import os

import numpy as np
import pytest

import i3elvis_amd as I
from conftest import golden

pytestmark = pytest.mark.gpu

ST = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def _scaled(a, b):
    """error relative to max(|b|, 1e-2 max|b|): sums whose terms change sign
    (the thermal expansion of molten rock) are judged on the field's scale"""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-2 * max(np.abs(b).max(), 1e-300))


def _oracle_sum64(g, densimod, **kw):
    """the oracle's ronurecalc with FP64 node sums (what the cell-segmented
    transfer computes, up to the order of the FP64 additions)"""
    import oracle
    oracle.set_sum64(True)
    try:
        fld, *_ = oracle.ronurecalc_x(
            g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["marke"], g["markrr"], g["markrn"], g["gx"],
            g["gy"], g["gz"], g["pr"], [g[k] for k in ST], [dict(zip(oracle.ROCKX_FIELDS, r)) for r in g["rocks"]],
            kw.pop("init"), float(g["nubeg"]), float(g["nuend"]), float(g["timestep"]), float(g["timesum"]),
            float(g["convend"]), nukoef=float(g["nukoef"]), densimod=densimod, vdeep=float(g["vdeep"]),
            nudeep=float(g["nudeep"]), eroslev=float(g["eroslev"]), sedilev=float(g["sedilev"]),
            waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]), mnumy=int(g["mnumy"]),
            teclmin=float(g["teclmin"]), teclmax=float(g["teclmax"]), **kw)
    finally:
        oracle.set_sum64(False)
    fld[6] = oracle.tkrecalc(fld[6], g["bondm"], g["bondv"], g["bondn"])
    return fld


def _ulp(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.mark.parametrize("case", [0, 1])
def test_movemark_rheology_matches_reference(gpu, case):
    g = golden("g14_movemark_rheology.npz")
    n = len(g["x"])
    print(f"case: movemark rheology g14 case {case} (outgrid {int(g['outgrid'][case])}, "
          f"markmod {int(g['markmod'][case])}), {n} markers", flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], n + 16)
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    M.set_strain(g["marke"])
    M.set_grain(g["markrr"], g["markrn"])
    rocks = [list(r) for r in g["rocks"]]
    dt = float(g["timestep"])
    M.update_props(rocks, float(g["nubeg"]), float(g["nuend"]), dt, float(g["healrate"]), [g[k] for k in ST],
                   g["pr"], rheology=1, timesum=float(g["timesum"][case]), convend=float(g["convend"]),
                   outgrid=int(g["outgrid"][case]))
    M.advect_ex(dt, g["vx"], g["vy"], g["vz"], outgrid=int(g["outgrid"][case]), markmod=int(g["markmod"][case]),
                vyfluid=float(g["vyfluid"]), vymelt=float(g["vymelt"]), g=tuple(g["g"]), pr=g["pr"],
                markd=g["markd"])
    d = M.download()
    e = M.get_strain()
    rr, rn = M.get_grain()
    assert np.array_equal(d["id"], np.arange(n))
    assert np.array_equal(d["type"], g[f"type_out{case}"])
    for name, a, tol in (("eta", d["eta"], 2e-6), ("rr", rr, 2e-6), ("rn", rn, 2e-6)):
        rel = _rel(a, g[f"{name}_out{case}"])
        print(f"  {name}: max rel {rel.max():.2e}, exact {int((a == g[f'{name}_out{case}']).sum())} of {n}", flush=True)
        assert rel.max() <= tol, (name, int(rel.argmax()))
    ref_e = g[f"e_out{case}"]
    assert np.array_equal(np.sign(e), np.sign(ref_e))
    assert np.abs(e - ref_e).max() <= 1e-6
    for k in ("x", "y", "z"):
        u = _ulp(d[k], g[f"{k}_out{case}"])
        print(f"  {k}: max ulp {u.max()}", flush=True)
        assert u.max() <= 2, (k, int(u.argmax()))
    assert np.array_equal(d["tk"] == 0, g[f"tk_out{case}"] == 0)
    M.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_ronurecalc_rheology_matches_reference(gpu, mode):
    """order 0: the reference's serial node sums; order 1: cell-segmented
    FP64 sums, within the reference's float accumulation rounding"""
    g = golden("g15_ronurecalc_rheology.npz")
    n = len(g["x"])
    print(f"case: ronurecalc rheology g15 order {mode}, {n} markers", flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], n + 16, order=mode)
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    M.set_strain(g["marke"])
    M.set_grain(g["markrr"], g["markrn"])
    M.set_stress([g[k] for k in ST])
    M.sort()
    M.set_tbc(g["bondm"], g["bondv"], g["bondn"])
    rocks = [list(r) for r in g["rocks"]]
    fld = M.to_nodes(rocks, g["pr"], g["init"], float(g["nubeg"]), float(g["nuend"]), nukoef=float(g["nukoef"]),
                     densimod=int(g["densimod"]), timestep=float(g["timestep"]), vdeep=float(g["vdeep"]),
                     nudeep=float(g["nudeep"]), rheology=1, eroslev=float(g["eroslev"]), sedilev=float(g["sedilev"]),
                     waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]), mnumy=int(g["mnumy"]),
                     teclmin=float(g["teclmin"]), teclmax=float(g["teclmax"]), timesum=float(g["timesum"]),
                     convend=float(g["convend"]))
    d = M.download()
    e = M.get_strain()
    rr, rn = M.get_grain()
    order = np.argsort(d["id"])
    assert np.array_equal(d["type"][order], g["type_out"])
    for name, a in (("eta", d["eta"]), ("rr", rr), ("rn", rn)):
        rel = _rel(a[order], g[f"{name}_out"])
        print(f"  {name}: max rel {rel.max():.2e}", flush=True)
        assert rel.max() <= 2e-6, name
    assert np.array_equal(np.sign(e[order]), np.sign(g["e_out"]))
    for q in range(11):
        rel = _rel(fld[q], g["out"][q]) if mode == 0 else _scaled(fld[q], g["out"][q])
        print(f"  field {q}: max rel {rel.max():.2e}", flush=True)
        assert rel.max() <= (2e-6 if mode == 0 else 4e-6), q
    if mode == 1:
        o = _oracle_sum64(g, int(g["densimod"]), init=g["init"])
        for q in range(11):
            assert _scaled(fld[q], o[q]).max() <= 2e-7, q
    M.close()


def _g15_cell_segmented(staged):
    """one cell-segmented ronurecalc of the g15 markers (whole rheology),
    the records staged in memory or not (I3G_MK_STAGED, read at create)"""
    g = golden("g15_ronurecalc_rheology.npz")
    n = len(g["x"])
    os.environ["I3G_MK_STAGED"] = "1" if staged else "0"
    try:
        M = I.Markers(g["gx"], g["gy"], g["gz"], n + 16, order=1)
    finally:
        del os.environ["I3G_MK_STAGED"]
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    M.set_strain(g["marke"])
    M.set_grain(g["markrr"], g["markrn"])
    M.set_stress([g[k] for k in ST])
    M.sort()
    M.set_tbc(g["bondm"], g["bondv"], g["bondn"])
    rocks = [list(r) for r in g["rocks"]]
    fld = M.to_nodes(rocks, g["pr"], g["init"], float(g["nubeg"]), float(g["nuend"]), nukoef=float(g["nukoef"]),
                     densimod=int(g["densimod"]), timestep=float(g["timestep"]), vdeep=float(g["vdeep"]),
                     nudeep=float(g["nudeep"]), rheology=1, eroslev=float(g["eroslev"]), sedilev=float(g["sedilev"]),
                     waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]), mnumy=int(g["mnumy"]),
                     teclmin=float(g["teclmin"]), teclmax=float(g["teclmax"]), timesum=float(g["timesum"]),
                     convend=float(g["convend"]))
    out = dict(fld=fld, e=M.get_strain(), grain=M.get_grain(), **M.download())
    M.close()
    return out


def test_fused_transfer_equals_staged_records(gpu):
    """The default one-pass transfer (each cell's eight corner lanes compute
    the properties of eight markers at a time and sum them from LDS) and the
    two-pass one that stages every marker's record in memory
    (I3G_MK_STAGED=1) give the same node fields and marker updates (type,
    viscosity, strain, grain) to the bit: the same records summed in the
    same order."""
    print("case: fused vs staged transfer, g15 whole rheology", flush=True)
    a, b = _g15_cell_segmented(False), _g15_cell_segmented(True)
    assert np.array_equal(a["fld"], b["fld"])
    for k in ("x", "y", "z", "tk", "eta", "type", "id", "e"):
        assert np.array_equal(a[k], b[k]), k
    for q in range(2):
        assert np.array_equal(a["grain"][q], b["grain"][q])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("densimod", [0, 1])
def test_ronurecalc_melt_matches_reference(gpu, densimod, mode):
    """partially molten rocks (melt types with water content and extracted
    melt, mark3mg.c:660-686): node fields, marker viscosities and types
    against the reference's ronurecalc (golden g17)"""
    g = golden("g17_ronurecalc_melt.npz")
    n = len(g["x"])
    print(f"case: ronurecalc melt g17 densimod {densimod} order {mode}, {n} markers", flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], n + 16, order=mode)
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    M.set_strain(g["marke"])
    M.set_grain(g["markrr"], g["markrn"])
    M.set_stress([g[k] for k in ST])
    M.set_melt(g["markw"], g["markex"])
    M.sort()
    M.set_tbc(g["bondm"], g["bondv"], g["bondn"])
    rocks = [list(r) for r in g["rocks"]]
    fld = M.to_nodes(rocks, g["pr"], np.zeros((11, len(g["pr"])), np.float32), float(g["nubeg"]),
                     float(g["nuend"]), nukoef=float(g["nukoef"]), densimod=densimod, timestep=float(g["timestep"]),
                     vdeep=float(g["vdeep"]), nudeep=float(g["nudeep"]), rheology=1, eroslev=float(g["eroslev"]),
                     sedilev=float(g["sedilev"]), waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]),
                     mnumy=int(g["mnumy"]), teclmin=float(g["teclmin"]), teclmax=float(g["teclmax"]),
                     timesum=float(g["timesum"]), convend=float(g["convend"]), maxpmelt=float(g["maxpmelt"]))
    d = M.download()
    order = np.argsort(d["id"])
    assert np.array_equal(d["type"][order], g[f"type_out{densimod}"])
    rel = _rel(d["eta"][order], g[f"eta_out{densimod}"])
    print(f"  eta: max rel {rel.max():.2e}, exact {int((d['eta'][order] == g[f'eta_out{densimod}']).sum())} of {n}",
          flush=True)
    assert rel.max() <= 2e-6
    for q in range(11):
        ref = g[f"out{densimod}"][q]
        rel = _rel(fld[q], ref) if mode == 0 else _scaled(fld[q], ref)
        print(f"  field {q}: max rel {rel.max():.2e}, exact {int((fld[q] == ref).sum())} of {ref.size}", flush=True)
        assert rel.max() <= (2e-6 if mode == 0 else 4e-6), q
    if mode == 1:
        o = _oracle_sum64(g, densimod, init=np.zeros((11, len(g["pr"])), np.float32),
                          maxpmelt=float(g["maxpmelt"]), markw=g["markw"], markex=g["markex"])
        for q in range(11):
            assert _scaled(fld[q], o[q]).max() <= 2e-7, q
    M.close()


def test_transfer_refuses_densimod_2(gpu):
    """the thermodynamic tables (densimod >= 2) are not restated"""
    g = golden("g15_ronurecalc_rheology.npz")
    n = len(g["x"])
    print("case: densimod 2 refused", flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], n + 16)
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    M.sort()
    rocks = [list(r) for r in g["rocks"]]
    with pytest.raises(I.I3GError):
        M.to_nodes(rocks, g["pr"], g["init"], 1e18, 1e25, densimod=2, rheology=1, mnumy=2)
    M.close()


def test_advect_rk4_equals_advect_ex_defaults(gpu):
    """i3g_markers_advect_rk4 is advect_ex with outgrid 0, markmod 0 and no
    fluid velocity: the same positions to the bit"""
    g = golden("g14_movemark_rheology.npz")
    n = len(g["x"])
    print("case: advect_rk4 vs advect_ex defaults", flush=True)
    out = []
    for ex in (False, True):
        M = I.Markers(g["gx"], g["gy"], g["gz"], n)
        M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
        if ex:
            M.advect_ex(3e11, g["vx"], g["vy"], g["vz"])
        else:
            M.advect_rk4(3e11, g["vx"], g["vy"], g["vz"])
        out.append(M.download())
        M.close()
    for k in ("x", "y", "z", "tk", "eta", "type"):
        assert np.array_equal(out[0][k], out[1][k]), k
