"""CPU tests: value-mode strain rates/stresses and the residual-mode rows
behind the reference's native norms, oracle against the reference's own
outputs (g7 fixture written from the compiled reference)."""
import numpy as np

import oracle
from conftest import golden


def _setup(g):
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    L = oracle.level(nx, ny, nz, 0, 0, g["gx"], g["gy"], g["gz"])
    et = oracle.eta(L, g["nu"], g["nxx"], g["nxy"], g["nxz"], g["nyz"], 0, float(g["nubeg"]), float(g["nuend"]))
    return L, et


def test_stress_matches_reference():
    g = golden("g7_stress_norms.npz")
    L, et = _setup(g)
    eps, sig = oracle.stress(L, et, g["vx"], g["vy"], g["vz"])
    ref = g["stress"]
    for q in range(6):
        for a, b in ((eps[q], ref[q]), (sig[q], ref[6 + q])):
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 2e-7 * scale, q


def test_err_rows_match_reference():
    g = golden("g7_stress_norms.npz")
    L, _ = _setup(g)
    N = len(g["nu"])
    sol = np.stack([g["pr"], g["vx"].astype(np.float64), g["vy"].astype(np.float64),
                    g["vz"].astype(np.float64)], 1).ravel()
    rows = oracle.err_rows(L, g["nu"], g["ro"], g["g"], float(g["nubeg"]), float(g["nuend"]), g["stress"], g["pr"],
                           g["vx"], g["vy"], g["vz"], g["bondm"], g["bondv"], g["bondn"], sol)
    ref = g["rows"]
    assert np.array_equal(np.isfinite(rows), np.isfinite(ref))
    f = np.isfinite(ref)
    assert np.allclose(rows[f], ref[f], rtol=1e-12, atol=0)
    # aggregated norms agree too
    a = oracle.native_norms(rows, 1e-9)
    b = oracle.native_norms(ref, 1e-9)
    for k in ("stoksum", "stoksum1", "stoksum2", "bondsum"):
        assert np.isclose(a[k], b[k], rtol=1e-12)
