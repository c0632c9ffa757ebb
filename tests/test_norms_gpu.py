"""GPU tests: strain rates/stresses and the reference's native norms of a
resident solution against the oracle (pinned to the reference by g7)."""
import numpy as np
import pytest

import i3elvis_amd as I
import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def test_stress_and_native_norms_match_oracle(gpu):
    g = golden("g7_stress_norms.npz")
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    N = nx * ny * nz
    print("case: g7 stress/norms", N, flush=True)
    cfg = dict(nx=nx, ny=ny, nz=nz, gx=g["gx"], gy=g["gy"], gz=g["gz"], multinum=0, bondm=g["bondm"],
               bondv=g["bondv"], bondn=g["bondn"], g=tuple(g["g"]), nubeg=float(g["nubeg"]),
               nuend=float(g["nuend"]), viscmod=0)
    for k in ("nu", "nxx", "nxy", "nxz", "nyz", "ro"):
        cfg[k] = g[k]
    S = I.Stokes(cfg)
    S.set_coefficients()
    S.set_solution(g["pr"], g["vx"].astype(np.float64), g["vy"].astype(np.float64), g["vz"].astype(np.float64))
    P, vx, vy, vz = S.get_solution()
    eps, sig = S.strain_stress()
    L = oracle.level(nx, ny, nz, 0, 0, g["gx"], g["gy"], g["gz"])
    et = oracle.eta(L, g["nu"], g["nxx"], g["nxy"], g["nxz"], g["nyz"], 0, float(g["nubeg"]), float(g["nuend"]))
    f32 = [a.astype(np.float32) for a in (vx, vy, vz)]
    e_o, s_o = oracle.stress(L, et, *f32)
    for q in range(6):
        for a, b in ((eps[q], e_o[q]), (sig[q], s_o[q])):
            scale = max(np.abs(b).max(), 1e-30)
            assert np.abs(a - b).max() <= 2e-7 * scale, q
    nn = S.native_norms()
    sol = np.stack([P, vx, vy, vz], 1).ravel()
    rows = oracle.err_rows(L, g["nu"], g["ro"], g["g"], float(g["nubeg"]), float(g["nuend"]),
                           np.concatenate([e_o, s_o]), P, *f32, g["bondm"], g["bondv"], g["bondn"], sol)
    mv = oracle.maxvxyz(vx, vy, vz, nx, ny, nz)
    ref = oracle.native_norms(rows, mv, g["bondm"][0::4][:N])
    print({k: (nn[k], ref.get(k)) for k in ("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum")},
          flush=True)
    assert np.isclose(nn["maxvxyz"], mv, rtol=1e-12)
    for k in ("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum"):
        assert np.isclose(nn[k], ref[k], rtol=1e-9, atol=1e-300), k
    S.close()


def test_native_norms_drop_on_converged_solve(gpu):
    n = 17
    cfg = I.make_config(0, n, n, n, seed=1)
    S = I.Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    S.solve(precision=64, rtol=1e-10)
    nn = S.native_norms()
    # the momentum residual of the converged fields, from float stresses,
    # is far below the buoyancy scale rho g
    assert nn["stoksum"] < 1e-3 * 3300 * 9.81
    assert nn["bondsum"] < 1e-12
    S.close()
