"""GPU parity tests of the Stokes path, through the C ABI (libi3g.so).

Tolerances: the FP64 residual kernel evaluates the same rows as the
reference in a different summation order, so it must agree with the oracle
and the reference's rows to 1e-12 of the row scale.  Converged fields must
be within 1e-6 relative L2 of the reference's converged FP64 solution
with the relative weighted residual <= 1e-8.
"""
import numpy as np
import pytest

import i3elvis_amd as I
import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _cfg_from_g1(g):
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    cfg = dict(nx=nx, ny=ny, nz=nz, gx=g["gx"], gy=g["gy"], gz=g["gz"], multinum=int(g["multinum"]),
               bondm=g["bondm"], bondv=g["bondv"], bondn=g["bondn"], nubeg=float(g["nubeg"]),
               nuend=float(g["nuend"]), g=tuple(g["g"]), viscmod=0)
    for k in ("nu", "nxx", "nxy", "nxz", "nyz", "ro"):
        cfg[k] = g[k]
    return cfg


def test_residual_fp64_matches_reference_rows(gpu):
    g = golden("g1_operator.npz")
    cfg = _cfg_from_g1(g)
    print("case: g1 residual", cfg["nx"], cfg["ny"], cfg["nz"], flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    x = g["x0"]
    r, norms = S.residual(x[0::4], x[1::4], x[2::4], x[3::4])
    L = oracle.level(cfg["nx"], cfg["ny"], cfg["nz"], 0, 0, cfg["gx"], cfg["gy"], cfg["gz"])
    et = oracle.eta(L, cfg["nu"], cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], 0, cfg["nubeg"], cfg["nuend"])
    r_or = oracle.residual(L, et, cfg["bondm"], cfg["bondv"], cfg["bondn"], g["rhs0"], x)
    ex = g["exists0"]
    for c in range(4):
        s = np.abs(g["r0"][c::4]).max() + np.abs(g["rhs0"][c::4]).max()
        d_ref = np.abs(r[c] - g["r0"][c::4])[ex[c::4]].max()
        d_or = np.abs(r[c] - r_or[c::4])[ex[c::4]].max()
        assert d_ref <= 1e-12 * s, (c, d_ref, s)
        assert d_or <= 1e-12 * s, (c, d_or, s)
    w = oracle.wnorms(L, et[0], cfg["bondm"], g["rhs0"], r_or)
    assert np.allclose(norms, w, rtol=1e-10)
    S.close()


@pytest.mark.parametrize("n", [17, 25])
@pytest.mark.parametrize("prec", [64, 32])
def test_solve_matches_reference_converged(gpu, n, prec):
    g = golden(f"g3_converged_c0_{n}.npz")
    cfg = I.make_config(0, n, n, n, seed=int(g["seed"]))
    print(f"case: solve c0 n={n} prec={prec} multinum={cfg['multinum']}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    N = n ** 3
    z = np.zeros(N)
    S.set_solution(z, z, z, z)
    st = S.solve(precision=prec, rtol=1e-9, max_outer=40)
    print("  outer", st.outer_iters, "vcycles", st.vcycles, "rel", st.rel_res, "s", st.seconds, flush=True)
    assert st.converged and st.rel_res <= 1e-9
    P, vx, vy, vz = S.get_solution()
    sol = g["sol"]
    ex = g["exists"]
    for c, f in enumerate((P, vx, vy, vz)):
        e = ex[c::4]
        ref = sol[c::4][e]
        err = np.linalg.norm(f[e] - ref) / np.linalg.norm(ref)
        print(f"  field {c}: rel L2 err {err:.3e}", flush=True)
        assert err <= 1e-6, (c, err)
    S.close()


def test_combine_start_recovers_solution(gpu):
    """i3g_stokes_combine_start: with the converged solution and a random
    vector as directions from a zero base, the residual-optimal coefficients
    are (1, 0) and the start is the solution."""
    import i3elvis_amd as I
    n = 17
    cfg = I.make_config(0, n, n, n, seed=1)
    S = I.Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    S.solve(precision=64, rtol=1e-11)
    xs = S.get_solution()
    rng = np.random.default_rng(3)
    rnd = [rng.standard_normal(n ** 3) * s for s in (1e6, 1e-10, 1e-10, 1e-10)]
    a, rel = S.combine_start([None] * 4, [xs, rnd])
    print("coefficients", a, "rel", rel, flush=True)
    assert abs(a[0] - 1) < 1e-6 and abs(a[1]) < 1e-6 * max(1.0, abs(a[0]))
    assert rel < 1e-9
    _, rel0 = S.combine_start([None] * 4, [])
    assert abs(rel0 - 1.0) < 1e-12
    S.close()


@pytest.mark.parametrize("prec,tol", [(64, 1e-11), (32, 1e-4)])
def test_load_first_smoother_matches_plain_kernel(gpu, monkeypatch, prec, tol):
    """The velocity smoother with all row loads issued first
    (k_smooth_vel_ilp, the default) and the plain row kernel
    (I3G_SMOOTH_ILP=0) compute the same update up to rounding: after the
    same V-cycles from a zero start the fields agree to rounding level."""
    n = 25
    cfg = I.make_config(2, n, n, n, seed=5)
    sols = []
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_SMOOTH_ILP", mode)
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=prec, rtol=1e-8, max_outer=1, max_vcycles=6, check_conv=False)
        print(f"case: smoother ILP={mode} prec={prec} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        d = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        print(f"  field {c}: rel L2 diff {d:.3e}", flush=True)
        assert d <= tol, (c, d)
