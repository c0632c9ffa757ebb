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


@pytest.mark.parametrize("which,n", [(0, 17), (0, 25), (2, 25), (2, 33)])
@pytest.mark.parametrize("prec", [64, 32])
def test_solve_matches_reference_converged(gpu, which, n, prec):
    """Converged fields against the direct solve of the reference's own
    assembled level-0 system (g3): configuration 0 (sphere) and the
    subduction layout of the 257^3 headline workload (10^6 contrast, weak
    channel) at 25^3 and 33^3, mixed precision and FP64, each field within
    1e-6 relative L2."""
    g = golden(f"g3_converged_c{which}_{n}.npz")
    cfg = I.make_config(which, n, n, n, seed=int(g["seed"]))
    print(f"case: solve c{which} n={n} prec={prec} multinum={cfg['multinum']}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    N = n ** 3
    z = np.zeros(N)
    S.set_solution(z, z, z, z)
    # the subduction system is ill-conditioned (10^6 contrast): a relative
    # residual of 5e-11 still leaves the pressure ~1e-6 off, so it is
    # solved to 1e-12
    rtol = 1e-9 if which == 0 else 1e-12
    st = S.solve(precision=prec, rtol=rtol, max_outer=40, max_vcycles=6000)
    print("  outer", st.outer_iters, "vcycles", st.vcycles, "rel", st.rel_res, "s", st.seconds, flush=True)
    assert st.converged and st.rel_res <= rtol
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


def test_solve_coarse_direct_matches_reference_converged(gpu, monkeypatch):
    """The coarsest level solved directly in FP64 (I3G_COARSE_DIRECT=1: the
    SVD of its assembled operator, singular values below 1e-17 of the
    largest dropped) inside the mixed V-cycle: the 33^3 subduction solve
    converges to the reference's direct solve (g3) within 1e-6 per field."""
    monkeypatch.setenv("I3G_COARSE_DIRECT", "1")
    monkeypatch.setenv("I3G_COARSE_DIRECT_TOL", "1e-17")
    g = golden("g3_converged_c2_33.npz")
    n = 33
    cfg = I.make_config(2, n, n, n, seed=int(g["seed"]))
    print(f"case: coarse direct solve n={n} multinum={cfg['multinum']}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    st = S.solve(precision=32, rtol=1e-12, max_outer=40, max_vcycles=6000)
    print("  outer", st.outer_iters, "vcycles", st.vcycles, "rel", st.rel_res, flush=True)
    assert st.converged and st.rel_res <= 1e-12
    sol, ex = g["sol"], g["exists"]
    for c, f in enumerate(S.get_solution()):
        e = ex[c::4]
        ref = sol[c::4][e]
        err = np.linalg.norm(f[e] - ref) / np.linalg.norm(ref)
        print(f"  field {c}: rel L2 err {err:.3e}", flush=True)
        assert err <= 1e-6, (c, err)
    S.close()


def test_scgs_smoother_converges_to_reference(gpu, monkeypatch):
    """The coupled cell smoother alone (I3G_SCGS=1 on every level, damping
    0.5, no colour phases, no guard) inside the mixed V-cycle: the 33^3
    subduction solve converges to the reference's direct solve (g3) within
    1e-6 per field, and reaches 1e-8 in fewer V-cycles than the colour
    phases alone (I3G_SCGS=0; 97 against 654 measured)."""
    g = golden("g3_converged_c2_33.npz")
    n = 33
    cfg = I.make_config(2, n, n, n, seed=int(g["seed"]))
    z = np.zeros(n ** 3)
    counts = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_SCGS", mode)
        monkeypatch.setenv("I3G_SCGS_LEVELS", "64")
        monkeypatch.setenv("I3G_SCGS_W", "0.5,0.5")
        monkeypatch.setenv("I3G_SCGS_MIX", "0")
        monkeypatch.setenv("I3G_SCGS_GUARD", "0")
        S = I.Stokes(cfg)
        S.set_coefficients()
        S.set_solution(z, z, z, z)
        st = S.solve(precision=32, rtol=1e-8, max_outer=40, max_vcycles=6000)
        counts[mode] = st.vcycles
        print(f"case: SCGS={mode} n={n}: {st.vcycles} V-cycles to rel {st.rel_res:.2e}", flush=True)
        if mode == "1":
            S.set_solution(z, z, z, z)
            st = S.solve(precision=32, rtol=1e-12, max_outer=40, max_vcycles=6000)
            assert st.converged and st.rel_res <= 1e-12
            sol, ex = g["sol"], g["exists"]
            for c, f in enumerate(S.get_solution()):
                e = ex[c::4]
                ref = sol[c::4][e]
                err = np.linalg.norm(f[e] - ref) / np.linalg.norm(ref)
                print(f"  field {c}: rel L2 err {err:.3e}", flush=True)
                assert err <= 1e-6, (c, err)
        S.close()
    assert counts["1"] < counts["0"], counts


@pytest.mark.parametrize("forced", [False, True])
def test_default_smoother_with_guard_converges_to_reference(gpu, monkeypatch, forced):
    """The default V-cycle smoother (level 0: at most six colour-phase
    sweeps, then one coupled cell sweep; damping 0.4/0.5) with the stall
    guard as it is (window 100 V-cycles, ratio 0.9: it does not fire here)
    and forced to fire (window 1, ratio 1e-3: the coupled sweeps are
    switched off after the first GCR direction, and the flexible GCR
    carries on with the full colour phases): the 33^3 subduction solve
    converges to the reference's direct solve (g3) within 1e-6 per field
    either way."""
    for k in ("I3G_SCGS", "I3G_SCGS_LEVELS", "I3G_SCGS_W", "I3G_SCGS_MIX", "I3G_SCGS_PHASES", "I3G_SCGS_GUARD",
              "I3G_SCGS_GUARD_RATIO"):
        monkeypatch.delenv(k, raising=False)
    if forced:
        monkeypatch.setenv("I3G_SCGS_GUARD", "1")
        monkeypatch.setenv("I3G_SCGS_GUARD_RATIO", "0.001")
    guard = "forced" if forced else "default"
    g = golden("g3_converged_c2_33.npz")
    n = 33
    cfg = I.make_config(2, n, n, n, seed=int(g["seed"]))
    S = I.Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    S.set_timing(True)
    st = S.solve(precision=32, rtol=1e-12, max_outer=40, max_vcycles=6000)
    calls = S.kernel_stats().get("scgs_L0", {}).get("calls", 0)
    print(f"case: default smoother, guard {guard}: {st.vcycles} V-cycles, {calls} coupled sweeps, "
          f"rel {st.rel_res:.2e}", flush=True)
    assert st.converged and st.rel_res <= 1e-12
    if forced:
        assert 0 < calls < 2 * st.vcycles  # switched off part-way
    else:
        assert calls == 2 * st.vcycles  # one coupled sweep before and after the coarse correction
    sol, ex = g["sol"], g["exists"]
    for c, f in enumerate(S.get_solution()):
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
    monkeypatch.setenv("I3G_FUSED_SWEEP", "0")  # the phase launches, where the two kernels differ
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


def _transfer_mask(exists, bondm, mgp, N, c):
    return exists[c::4] & (bondm[4 * mgp:4 * (mgp + N)][c::4] == 0)


def test_transfers_match_reference_bitwise(gpu):
    """i3g_stokes_restrict / i3g_stokes_prolong (FP64) against vprestrict /
    vpprolong run by the reference on the same inputs (g1 fixture, 13x11x12
    -> 9x7x8): equal to the bit on every existing non-BC coarse row and on
    every existing fine unknown; the other coarse rows are 0 and missing
    fine unknowns are untouched."""
    g = golden("g1_operator.npz")
    cfg = _cfg_from_g1(g)
    print("case: g1 transfers", cfg["nx"], cfg["ny"], cfg["nz"], flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    rr = g["restr_in"]
    out = S.restrict(1, [rr[c::4] for c in range(4)])
    p1, N1 = int(g["mgp"][1]), len(out[0])
    for c in range(4):
        m = _transfer_mask(g["exists1"], g["bondm"], p1, N1, c)
        ref = g["restr_out"][c::4]
        print(f"  restrict comp {c}: rows {int(m.sum())} max|diff| {np.abs(out[c][m] - ref[m]).max():.3e}", flush=True)
        assert np.array_equal(out[c][m], ref[m]), c
        assert np.all(out[c][~m] == 0), c
    xf, xc = g["prol_xf"], g["prol_xc"]
    pf = S.prolong(1, [xc[c::4] for c in range(4)], [xf[c::4] for c in range(4)], float(g["v1koef"]),
                   float(g["p1koef"]))
    for c in range(4):
        m = g["exists0"][c::4]
        ref = g["prol_out"][c::4]
        print(f"  prolong comp {c}: max|diff| {np.abs(pf[c][m] - ref[m]).max():.3e}", flush=True)
        assert np.array_equal(pf[c][m], ref[m]), c
        assert np.array_equal(pf[c][~m], xf[c::4][~m]), c
    S.close()


@pytest.mark.parametrize("vm", [1, 2])
def test_residual_viscmod_matches_reference_rows(gpu, vm):
    """Level-0 FP64 residual with geometric (1) / harmonic (2) viscosity
    means against the reference's rows (g9) to 1e-12 of the row scale."""
    g = golden("g1_operator.npz")
    r = golden("g9_viscmod.npz")
    cfg = _cfg_from_g1(g)
    cfg["viscmod"] = vm
    print(f"case: g9 residual viscmod={vm}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    x = r[f"x0_{vm}"]
    res, _ = S.residual(x[0::4], x[1::4], x[2::4], x[3::4])
    ex = r[f"exists0_{vm}"]
    for c in range(4):
        ref = r[f"r0_{vm}"][c::4]
        s = np.abs(ref).max() + np.abs(r[f"rhs0_{vm}"][c::4]).max()
        d = np.abs(res[c] - ref)[ex[c::4]].max()
        assert d <= 1e-12 * s, (vm, c, d, s)
    S.close()


@pytest.mark.parametrize("vm", [1, 2])
def test_solve_viscmod_satisfies_oracle_system(gpu, vm):
    """A converged solve with viscmod 1 / 2 (level-0 means and the coarse
    levels' ronurestrict) is a solution of the oracle's system with the
    same means: its oracle-weighted relative residual is <= 1e-8."""
    n = 25
    cfg = I.make_config(1, n, n, n, seed=4)
    cfg["viscmod"] = vm
    print(f"case: solve c1 n={n} viscmod={vm}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    z = np.zeros(n ** 3)
    S.set_solution(z, z, z, z)
    st = S.solve(precision=32, rtol=1e-9, max_outer=40)
    print("  vcycles", st.vcycles, "rel", st.rel_res, flush=True)
    assert st.converged
    P, vx, vy, vz = S.get_solution()
    S.close()
    L = oracle.level(n, n, n, 0, 0, cfg["gx"], cfg["gy"], cfg["gz"])
    et = oracle.eta(L, cfg["nu"], cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], vm, cfg["nubeg"], cfg["nuend"])
    b = oracle.rhs(L, cfg["ro"], cfg["g"], cfg["bondm"], cfg["bondv"])
    x = np.zeros(4 * n ** 3)
    for c, f in enumerate((P, vx, vy, vz)):
        x[c::4] = f
    res = oracle.residual(L, et, cfg["bondm"], cfg["bondv"], cfg["bondn"], b, x)
    w = oracle.wnorms(L, et[0], cfg["bondm"], b, res)
    print("  oracle rel residual", w[0] / w[1], flush=True)
    assert w[0] / w[1] <= 1e-8
    # and it is not a solution of the arithmetic-mean system
    et0 = oracle.eta(L, cfg["nu"], cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], 0, cfg["nubeg"], cfg["nuend"])
    res0 = oracle.residual(L, et0, cfg["bondm"], cfg["bondv"], cfg["bondn"], b, x)
    w0 = oracle.wnorms(L, et0[0], cfg["bondm"], b, res0)
    assert w0[0] / w0[1] > 1e-6


def test_transfers_match_oracle_all_levels(gpu):
    """Restriction and prolongation between every pair of levels of a 129^3
    hierarchy (129 -> 67 -> 36 -> 20 -> 12; the 67 -> 36 maps have
    five-line inverse lists) against the oracle's vprestrict / vpprolong,
    with the oracle's own coarse viscosities (ronurestrict) for the P
    scaling: bit-equal on existing non-BC coarse rows and existing fine
    unknowns."""
    n = 129
    cfg = I.make_config(2, n, n, n, seed=3)
    h = cfg["hier"]
    print("case: transfers c2 n=129 levels", cfg["multinum"] + 1, flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    rng = np.random.default_rng(11)
    Ls, nds = [], []
    nu = cfg["nu"]
    for l in range(cfg["multinum"] + 1):
        lines = I.level_lines(h, l, cfg["gx"], cfg["gy"], cfg["gz"])
        L = oracle.level(int(h.mgx[l]), int(h.mgy[l]), int(h.mgz[l]), int(h.mgp[l]), l, *lines)
        if l == 0:
            et = oracle.eta(L, nu, cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], 0, cfg["nubeg"], cfg["nuend"])
        else:
            nu = oracle.ronurestrict(Ls[-1], L, nu, 0)
            et = oracle.eta(L, nu, None, None, None, None, 0, cfg["nubeg"], cfg["nuend"])
        Ls.append(L)
        nds.append(et[0].astype(np.float32))
    for l in range(1, cfg["multinum"] + 1):
        F, Cl = Ls[l - 1], Ls[l]
        Nf, Nc = F.nx * F.ny * F.nz, Cl.nx * Cl.ny * Cl.nz
        rf = rng.standard_normal(4 * Nf)
        out = S.restrict(l, [rf[c::4] for c in range(4)])
        ref = oracle.vprestrict(F, Cl, cfg["bondm"], rf, nds[l - 1], nds[l])
        assert S.level_nodes(l) == Nc
        exc = np.zeros(4 * Nc, bool)
        mz, rem = np.divmod(np.arange(Nc), Cl.nx * Cl.ny)
        mx, my = np.divmod(rem, Cl.ny)
        exc[0::4] = (mx > 0) & (my > 0) & (mz > 0)
        exc[1::4] = (my < Cl.ny - 1) & (mz < Cl.nz - 1)
        exc[2::4] = (mx < Cl.nx - 1) & (mz < Cl.nz - 1)
        exc[3::4] = (mx < Cl.nx - 1) & (my < Cl.ny - 1)
        for c in range(4):
            m = _transfer_mask(exc, cfg["bondm"], int(h.mgp[l]), Nc, c)
            d = np.abs(out[c][m] - ref[c::4][m]).max()
            print(f"  level {l} restrict comp {c}: max|diff| {d:.3e}", flush=True)
            assert np.array_equal(out[c][m], ref[c::4][m]), (l, c)
        xc = rng.standard_normal(4 * Nc)
        xf = rng.standard_normal(4 * Nf)
        pf = S.prolong(l, [xc[c::4] for c in range(4)], [xf[c::4] for c in range(4)], 0.7, 1.0)
        pref = oracle.vpprolong(F, Cl, xc, xf, 0.7, 1.0)
        for c in range(4):
            d = np.abs(pf[c] - pref[c::4]).max()
            print(f"  level {l} prolong comp {c}: max|diff| {d:.3e}", flush=True)
            assert np.array_equal(pf[c], pref[c::4]), (l, c)
    S.close()


def test_create_rejects_degenerate_viscosity_restriction(gpu):
    """On a 68^3 grid the reference's coarse lines (68 -> 36) give a coarse
    line whose ronurestrict weights sum to zero (fine weights -1, 0, 1, 0),
    i.e. a division by zero in the coarse viscosity; the context refuses
    such a hierarchy instead of producing non-finite viscosities."""
    cfg = I.make_config(2, 68, 68, 68, seed=3)
    with pytest.raises(I.I3GError, match="invalid"):
        I.Stokes(cfg)


@pytest.mark.parametrize("n", [33, 65])
@pytest.mark.parametrize("prec", [32, 64])
def test_fused_bottom_matches_launched_levels(gpu, monkeypatch, n, prec):
    """k_coarsest (all sweeps of the coarsest level in one work-group, from
    LDS) against the same sweeps as separate launches
    (I3G_FUSED_BOTTOM=0): the same per-node code in the same pass order,
    so the fields after a few V-cycles are equal to the bit (coarsest
    level 12^3 for both sizes)."""
    cfg = I.make_config(2, n, n, n, seed=7)
    sols = []
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_FUSED_BOTTOM", mode)
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=prec, rtol=1e-8, max_outer=1, max_vcycles=5, check_conv=False)
        print(f"case: fused bottom={mode} n={n} prec={prec} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        assert np.array_equal(a, b), (c, np.abs(a - b).max())


@pytest.mark.parametrize("prec", [32, 64])
def test_bc_chain_matches_depth_launches(gpu, monkeypatch, prec):
    """BC rows of all dependency depths in one launch (k_bc_chain, partner
    chains evaluated innermost first) against one launch per depth
    (I3G_BC_CHAIN=0): equal to the bit after a few V-cycles, and for the
    inhomogeneous rows applied by set_solution."""
    n = 33
    cfg = I.make_config(2, n, n, n, seed=9)
    sols = []
    rng = np.random.default_rng(5)
    guess = [rng.standard_normal(n ** 3) * s for s in (1e6, 1e-10, 1e-10, 1e-10)]
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_BC_CHAIN", mode)
        S = I.Stokes(cfg)
        S.set_coefficients()
        S.set_solution(*guess)
        g = S.get_solution()
        st = S.solve(precision=prec, rtol=1e-8, max_outer=1, max_vcycles=4, check_conv=False)
        print(f"case: bc chain={mode} prec={prec} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append((g, S.get_solution()))
        S.close()
    for (ga, a), (gb, b) in [(sols[0], sols[1])]:
        for c in range(4):
            assert np.array_equal(ga[c], gb[c]), ("guess", c)
            assert np.array_equal(a[c], b[c]), ("solve", c)
