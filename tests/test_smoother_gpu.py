"""GPU parity of the V-cycle smoother, through the C ABI (libi3g.so).

One colour phase of the product's kernels against gausmat4's update rule
(move3mg.c:3100-3128) evaluated from the reference's own assembled rows
(golden g1: r0 = b - A x0; g12: each row's own coefficient a_ii and nd):
velocity x += v0koef r / a_ii on one colour of one component, pressure
P += p0koef nd r on every continuity row, everything else unchanged.  FP64
to 1e-10 of the phase's largest increment (the rows are summed in another
order than the reference's), FP32 to 1e-4 (the residual of a random state
in float).  The single-pass sweep (one launch per velocity component, out
of place), the z-slab schedule of the seven phases and the one-launch
sweep (k_sweep7: (x, y) tiles marching through the planes, phases lagging
0,1,1,2,2,3,3 planes, halo lines recomputed) and the three-launch sweep
(k_sweep3: one launch per velocity component, both colours through LDS)
against the seven phase launches: equal to the bit.
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


def _planes(v):
    return [v[c::4].copy() for c in range(4)]


def _phase_rows(nx, ny, nz, phase):
    k, i, j = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    n = (k * nx * ny + i * ny + j).ravel()
    if phase == 6:
        return n * 4
    sel = ((i + j + k).ravel() & 1) == (phase & 1)
    return n[sel] * 4 + 1 + phase // 2


@pytest.mark.parametrize("prec,tol", [(64, 1e-10), (32, 1e-4)])
@pytest.mark.parametrize("phase", range(7))
def test_smoother_phase_matches_reference_rule(gpu, phase, prec, tol):
    g = golden("g1_operator.npz")
    s = golden("g12_smoother.npz")
    cfg = _cfg_from_g1(g)
    print(f"case: smoother phase {phase} prec {prec} on g1 level 0", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    v0, p0 = float(s["v0koef"]), float(s["p0koef"])
    x0, r0 = g["x0"], g["r0"]
    out = S.smooth_phase(0, phase, prec, _planes(x0), _planes(g["rhs0"]), v0, p0)
    S.close()
    got = np.stack(out, 1).ravel()
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    bc = g["bondm"][:4 * nx * ny * nz] != 0
    rows = _phase_rows(nx, ny, nz, phase)
    rows = rows[g["exists0"][rows] & ~bc[rows]]
    if phase == 6:
        want = p0 * s["nd0"][rows // 4].astype(np.float64) * r0[rows]
    else:
        want = v0 * r0[rows] / s["diag0"][rows]
    xin = x0 if prec == 64 else x0.astype(np.float32).astype(np.float64)
    d = np.abs((got[rows] - xin[rows]) - want).max()
    print(f"  max |increment - reference rule| {d:.3e} of {np.abs(want).max():.3e}", flush=True)
    assert d <= tol * np.abs(want).max()
    others = np.ones(len(x0), bool)
    others[rows] = False
    assert np.array_equal(got[others], xin[others])


@pytest.mark.parametrize("phase", range(7))
def test_smoother_phase_matches_oracle_level1(gpu, phase):
    """The same phases on the coarse level of g1 (viscosity restricted by
    ronurestrict, pinned bitwise to the reference) against the oracle's
    or_colour_phase, FP64."""
    g = golden("g1_operator.npz")
    cfg = _cfg_from_g1(g)
    print(f"case: smoother phase {phase} on g1 level 1", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    x1, b1 = g["x1"], g["b1"]
    out = S.smooth_phase(1, phase, 64, _planes(x1), _planes(b1), 0.95, 0.7)
    S.close()
    L0 = oracle.level(cfg["nx"], cfg["ny"], cfg["nz"], 0, 0, g["gx"], g["gy"], g["gz"])
    L1 = oracle.level(int(g["mgx"][1]), int(g["mgy"][1]), int(g["mgz"][1]), int(g["mgp"][1]), 1,
                      g["lx1"], g["ly1"], g["lz1"])
    nu1 = oracle.ronurestrict(L0, L1, g["nu"], 0)
    et1 = oracle.eta(L1, nu1, None, None, None, None, 0, float(g["nubeg"]), float(g["nuend"]))
    ref = oracle.colour_phase(L1, et1, g["bondm"], b1, x1, phase, 0.7, 0.95)
    got = np.stack(out, 1).ravel()
    inc_ref, inc = ref - x1, got - x1
    d = np.abs(inc - inc_ref).max()
    print(f"  max |increment diff| {d:.3e} of {np.abs(inc_ref).max():.3e}", flush=True)
    assert d <= 1e-10 * np.abs(inc_ref).max()
    assert np.array_equal(got[inc_ref == 0], x1[inc_ref == 0])


def _random_state(N, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(N) * s for s in (1e6, 1e-10, 1e-10, 1e-10)]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case", ["g1_tiles_4_4_3", "g1_default", "sub65_default", "sub65_tiles_5_7_3",
                                  "sub65_level1"])
def test_single_pass_sweep_matches_phase_launches(gpu, monkeypatch, prec, case):
    """k_sweep_comp (variant 2: three launches a sweep, out of place, tiles with
    recomputed halo rings, z-slabs) equals the seven in-place phase launches
    plus BC rows to the bit, for 2 and 4 sweeps (pairs: the result returns to
    the level's vector) on grids whose tiles and slabs do not divide evenly."""
    if case.startswith("g1"):
        g = golden("g1_operator.npz")
        cfg = _cfg_from_g1(g)
        level = 0
    else:
        cfg = I.make_config(2, 65, 65, 65, seed=3)
        level = 1 if case.endswith("level1") else 0
    monkeypatch.setenv("I3G_FUSED_MINZ", "0")  # every level, however thin
    if "tiles" in case:
        monkeypatch.setenv("I3G_FUSED_TILE", ",".join(case.split("_")[-3:]))
    else:
        monkeypatch.delenv("I3G_FUSED_TILE", raising=False)
    print(f"case: single-pass sweep {case} prec {prec} level {level}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    Nl = S.level_nodes(level)
    x = _random_state(Nl, 11)
    b = _random_state(Nl, 12)
    for ns in (2, 4):
        a = S.smooth(level, prec, ns, x, b, 0.95, 0.7, variant=2)
        r = S.smooth(level, prec, ns, x, b, 0.95, 0.7, variant=1)
        for c in range(4):
            assert np.array_equal(a[c], r[c]), (ns, c, np.abs(a[c] - r[c]).max())
        assert any(not np.array_equal(a[c], x[c]) for c in range(4))
    S.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_single_pass_vcycles_match_phase_launch_vcycles(gpu, monkeypatch, prec):
    """Whole V-cycles (all levels, graph-captured coarse part) with the
    single-pass sweeps (I3G_FUSED_SWEEP=1) against the phase launches: the
    fields after a few V-cycles are equal to the bit."""
    n = 33
    cfg = I.make_config(2, n, n, n, seed=7)
    sols = []
    monkeypatch.setenv("I3G_FUSED_MINZ", "0")
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_FUSED_SWEEP", mode)
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=prec, rtol=1e-8, max_outer=1, max_vcycles=5, check_conv=False)
        print(f"case: fused sweep={mode} prec={prec} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        assert np.array_equal(a, b), (c, np.abs(a - b).max())


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("case,slab", [("g1", 1), ("g1", 2), ("sub65", 1), ("sub65", 3), ("sub65", 16),
                                       ("sub65", 50), ("sub65_level1", 2)])
def test_slab_sweep_matches_phase_launches(gpu, monkeypatch, prec, case, slab):
    """The seven phases in z-slabs (variant 3: phase q of slab [b, b') on
    planes [b - q, b' - q)) equal the phase launches over whole levels to the
    bit, for slabs thinner than the seven-plane lag, slabs that do not divide
    the level, and one sweep and three."""
    if case == "g1":
        cfg = _cfg_from_g1(golden("g1_operator.npz"))
        level = 0
    else:
        cfg = I.make_config(2, 65, 65, 65, seed=3)
        level = 1 if case.endswith("level1") else 0
    monkeypatch.setenv("I3G_SLAB", str(slab))
    print(f"case: slab sweep {case} slab {slab} prec {prec} level {level}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    Nl = S.level_nodes(level)
    x = _random_state(Nl, 21)
    b = _random_state(Nl, 22)
    for ns in (1, 3):
        a = S.smooth(level, prec, ns, x, b, 0.95, 0.7, variant=3)
        r = S.smooth(level, prec, ns, x, b, 0.95, 0.7, variant=1)
        for c in range(4):
            assert np.array_equal(a[c], r[c]), (ns, c, np.abs(a[c] - r[c]).max())
        assert any(not np.array_equal(a[c], x[c]) for c in range(4))
    S.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_slab_vcycles_match_phase_launch_vcycles(gpu, monkeypatch, prec):
    """Whole V-cycles with level-0 sweeps in z-slabs (I3G_SLAB=5) against the
    phase launches: the fields after a few V-cycles are equal to the bit."""
    n = 33
    cfg = I.make_config(2, n, n, n, seed=7)
    sols = []
    for slab in ("0", "5"):
        monkeypatch.setenv("I3G_SLAB", slab)
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=prec, rtol=1e-8, max_outer=1, max_vcycles=5, check_conv=False)
        print(f"case: slab={slab} prec={prec} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        assert np.array_equal(a, b), (c, np.abs(a - b).max())


@pytest.mark.parametrize("case", ["g1", "sub65", "sub65_level1", "sub_2x3", "c4_49x45x21"])
def test_sweep7_matches_phase_launches(gpu, case):
    """k_sweep7 (variant 4: one launch a FP32 sweep, out of place in pairs)
    equals the seven phase launches plus BC rows to the bit: one tile
    covering the whole plane (g1, 13x11x12, and the 36^3 level 1), 4x4
    uneven tiles (65^3), tiles of 16/17 lines on a non-cubic grid with few
    planes, and 2, 4 and 3 sweeps (an odd last sweep runs as phases)."""
    if case == "g1":
        cfg = _cfg_from_g1(golden("g1_operator.npz"))
        level = 0
    elif case.startswith("sub65"):
        cfg = I.make_config(2, 65, 65, 65, seed=3)
        level = 1 if case.endswith("level1") else 0
    elif case == "sub_2x3":
        cfg = I.make_config(2, 33, 51, 17, seed=5)
        level = 0
    else:
        cfg = I.make_config(2, 49, 45, 21, seed=6)
        level = 0
    print(f"case: sweep7 {case} level {level}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    Nl = S.level_nodes(level)
    x = _random_state(Nl, 31)
    b = _random_state(Nl, 32)
    for ns in (2, 4, 3):
        a = S.smooth(level, 32, ns, x, b, 0.95, 0.7, variant=4)
        r = S.smooth(level, 32, ns, x, b, 0.95, 0.7, variant=1)
        for c in range(4):
            assert np.array_equal(a[c], r[c]), (ns, c, np.abs(a[c] - r[c]).max())
        assert any(not np.array_equal(a[c], x[c]) for c in range(4))
    S.close()


def test_sweep7_vcycles_match_phase_launch_vcycles(gpu, monkeypatch):
    """Whole mixed-precision V-cycles with k_sweep7 on every level
    (I3G_SWEEP7=1, I3G_SWEEP7_MIN_TILES=1) against the phase launches
    (I3G_SWEEP7=0): the fields after a few V-cycles are equal to the bit."""
    n = 41
    cfg = I.make_config(2, n, n, n, seed=7)
    sols = []
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_SWEEP7", mode)
        monkeypatch.setenv("I3G_SWEEP7_MIN_TILES", "1")
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=32, rtol=1e-8, max_outer=1, max_vcycles=5, check_conv=False)
        print(f"case: sweep7={mode} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        assert np.array_equal(a, b), (c, np.abs(a - b).max())


@pytest.mark.parametrize("case,nzs", [("g1", 1), ("g1", 3), ("sub65", 1), ("sub65", 2), ("sub65_level1", 2),
                                      ("sub_2x3", 1), ("c4_49x45x21", 2), ("c4_49x45x21", 21)])
def test_sweep3_matches_phase_launches(gpu, monkeypatch, case, nzs):
    """k_sweep3 (variant 5: three launches a FP32 sweep, out of place in
    pairs) equals the seven phase launches plus BC rows to the bit: one tile
    over the whole plane (g1, the 35^3 level 1), uneven tiles of 24 x 28
    lines (65^3, 49x45x21), z-slabs of 1, 3, 12 or more planes, and 2, 4
    and 3 sweeps (an odd last sweep runs as phases)."""
    if case == "g1":
        cfg = _cfg_from_g1(golden("g1_operator.npz"))
        level = 0
    elif case.startswith("sub65"):
        cfg = I.make_config(2, 65, 65, 65, seed=3)
        level = 1 if case.endswith("level1") else 0
    elif case == "sub_2x3":
        cfg = I.make_config(2, 33, 51, 17, seed=5)
        level = 0
    else:
        cfg = I.make_config(2, 49, 45, 21, seed=6)
        level = 0
    monkeypatch.setenv("I3G_SWEEP3_NZS", str(nzs))
    print(f"case: sweep3 {case} level {level} nzs {nzs}", flush=True)
    S = I.Stokes(cfg)
    S.set_coefficients()
    Nl = S.level_nodes(level)
    x = _random_state(Nl, 41)
    b = _random_state(Nl, 42)
    for ns in (2, 4, 3):
        a = S.smooth(level, 32, ns, x, b, 0.95, 0.7, variant=5)
        r = S.smooth(level, 32, ns, x, b, 0.95, 0.7, variant=1)
        for c in range(4):
            assert np.array_equal(a[c], r[c]), (ns, c, np.abs(a[c] - r[c]).max())
        assert any(not np.array_equal(a[c], x[c]) for c in range(4))
    S.close()


def test_sweep3_vcycles_match_phase_launch_vcycles(gpu, monkeypatch):
    """Whole mixed-precision V-cycles with k_sweep3 on every level
    (I3G_SWEEP3_MIN_GROUPS=1) against the phase launches (I3G_SWEEP3=0): the
    fields after a few V-cycles are equal to the bit."""
    n = 41
    cfg = I.make_config(2, n, n, n, seed=7)
    sols = []
    for mode in ("0", "1"):
        monkeypatch.setenv("I3G_SWEEP3", mode)
        monkeypatch.setenv("I3G_SWEEP3_MIN_GROUPS", "1")
        S = I.Stokes(cfg)
        S.set_coefficients()
        z = np.zeros(n ** 3)
        S.set_solution(z, z, z, z)
        st = S.solve(precision=32, rtol=1e-8, max_outer=1, max_vcycles=5, check_conv=False)
        print(f"case: sweep3={mode} vcycles={st.vcycles} rel={st.rel_res:.3e}", flush=True)
        sols.append(S.get_solution())
        S.close()
    for c, (a, b) in enumerate(zip(*sols)):
        assert np.array_equal(a, b), (c, np.abs(a - b).max())
