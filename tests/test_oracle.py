"""CPU tests: the oracle (C restatement) against the reference's own outputs
(golden fixtures written from the compiled reference) and against closed
forms that a plausible mistake would break."""
import numpy as np
import pytest

import oracle
from conftest import golden


def _levels(g):
    L0 = oracle.level(int(g["nx"]), int(g["ny"]), int(g["nz"]), 0, 0, g["gx"], g["gy"], g["gz"])
    L1 = oracle.level(int(g["mgx"][1]), int(g["mgy"][1]), int(g["mgz"][1]), int(g["mgp"][1]), 1,
                      g["lx1"], g["ly1"], g["lz1"])
    return L0, L1


def _etas(g, L, lev, nu):
    if lev == 0:
        return oracle.eta(L, nu, g["nxx"], g["nxy"], g["nxz"], g["nyz"], 0, float(g["nubeg"]), float(g["nuend"]))
    return oracle.eta(L, nu, None, None, None, None, 0, float(g["nubeg"]), float(g["nuend"]))


def _rowscale(x, L, etas):
    """|A| |x| bound per unknown for a relative tolerance."""
    return np.abs(x).max() * 0 + 1.0


def test_rhs_matches_reference():
    g = golden("g1_operator.npz")
    L0, _ = _levels(g)
    b = oracle.rhs(L0, g["ro"], g["g"], g["bondm"], g["bondv"])
    assert np.array_equal(b, g["rhs0"])


@pytest.mark.parametrize("lev", [0, 1])
def test_residual_matches_reference_rows(lev):
    g = golden("g1_operator.npz")
    L0, L1 = _levels(g)
    if lev == 0:
        L, nu, x, b, rref, ex = L0, g["nu"], g["x0"], g["rhs0"], g["r0"], g["exists0"]
    else:
        L, nu, x, b, rref, ex = L1, g["nu1"], g["x1"], g["b1"], g["r1"], g["exists1"]
    et = _etas(g, L, lev, nu)
    r = oracle.residual(L, et, g["bondm"], g["bondv"], g["bondn"], b, x)
    # scale: the largest single term of a row ~ max|coef| * max|x| per component
    for c in range(4):
        d = np.abs(r[c::4] - rref[c::4])[ex[c::4]]
        s = np.abs(rref[c::4]).max() + np.abs(b[c::4]).max()
        assert d.max() <= 1e-12 * max(s, 1e-300) or d.max() < 1e-30, (c, d.max(), s)
    assert np.all(r[~ex] == 0)


def test_ronurestrict_bitwise():
    g = golden("g1_operator.npz")
    L0, L1 = _levels(g)
    nu1 = oracle.ronurestrict(L0, L1, g["nu"], 0)
    assert np.array_equal(nu1, g["nu1"])


def test_vprestrict_bitwise():
    g = golden("g1_operator.npz")
    L0, L1 = _levels(g)
    out = oracle.vprestrict(L0, L1, g["bondm"], g["restr_in"], g["nd0"], g["nd1"])
    assert np.array_equal(out, g["restr_out"])


def test_vpprolong_bitwise():
    g = golden("g1_operator.npz")
    L0, L1 = _levels(g)
    out = oracle.vpprolong(L0, L1, g["prol_xc"], g["prol_xf"], float(g["v1koef"]), float(g["p1koef"]))
    assert np.array_equal(out, g["prol_out"])


def test_nd_is_float_cell_viscosity():
    g = golden("g1_operator.npz")
    L0, _ = _levels(g)
    et = _etas(g, L0, 0, g["nu"])
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    bc = g["bondm"][:4 * nx * ny * nz:4] != 0
    n = np.arange(nx * ny * nz)
    m3, rem = np.divmod(n, nx * ny)
    m1, m2 = np.divmod(rem, ny)
    sel = (m1 > 0) & (m2 > 0) & (m3 > 0) & ~bc
    assert np.array_equal(et[0][sel].astype(np.float32), g["nd0"][sel])


def test_constant_viscosity_hydrostatic_closed_form():
    """Uniform eta and rho in a free-slip box: v = 0 and P = rho*g*depth
    (cell-centre depth) satisfy every interior row exactly."""
    import i3elvis_amd as I
    n = 11
    cfg = I.make_config(0, n, n, n, multinum=0)
    N = n ** 3
    for k in ("nu", "nxx", "nxy", "nxz", "nyz"):
        cfg[k][:] = 1e21
    cfg["ro"][:] = 3300.0
    L = oracle.level(n, n, n, 0, 0, cfg["gx"], cfg["gy"], cfg["gz"])
    et = oracle.eta(L, cfg["nu"], cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], 0, 1e18, 1e25)
    b = oracle.rhs(L, cfg["ro"], (0.0, 9.81, 0.0), cfg["bondm"], cfg["bondv"])
    x = np.zeros(4 * N)
    ymid = np.r_[0.0, 0.5 * (cfg["gy"][1:] + cfg["gy"][:-1])]
    j = np.arange(N) % n
    x[0::4] = 3300.0 * 9.81 * ymid[j]
    r = oracle.residual(L, et, cfg["bondm"], cfg["bondv"], cfg["bondn"], b, x)
    bc = cfg["bondm"][:4 * N] != 0
    assert np.abs(r[~bc]).max() < 1e-9 * 3300 * 9.81
    # a wrong sign on the pressure gradient breaks it
    r2 = oracle.residual(L, et, cfg["bondm"], cfg["bondv"], cfg["bondn"], b, -x)
    assert np.abs(r2[~bc]).max() > 1.0


@pytest.mark.parametrize("which,n", [(0, 17), (0, 25), (2, 25), (2, 33)])
def test_golden_solution_satisfies_oracle_system(which, n):
    """The reference-converged solution zeroes the oracle's residual
    (configuration 0 and the subduction layout)."""
    import i3elvis_amd as I
    g = golden(f"g3_converged_c{which}_{n}.npz")
    cfg = I.make_config(which, n, n, n, seed=int(g["seed"]))
    L = oracle.level(n, n, n, 0, 0, cfg["gx"], cfg["gy"], cfg["gz"])
    et = oracle.eta(L, cfg["nu"], cfg["nxx"], cfg["nxy"], cfg["nxz"], cfg["nyz"], 0, cfg["nubeg"], cfg["nuend"])
    b = oracle.rhs(L, cfg["ro"], cfg["g"], cfg["bondm"], cfg["bondv"])
    assert np.array_equal(b, g["rhs"])
    r = oracle.residual(L, et, cfg["bondm"], cfg["bondv"], cfg["bondn"], b, g["sol"])
    w = oracle.wnorms(L, et[0], cfg["bondm"], b, r)
    assert w[0] / w[1] < 1e-10


@pytest.mark.parametrize("vm", [1, 2])
def test_viscmod_rows_and_restriction_match_reference(vm):
    """viscmod 1 (geometric) and 2 (harmonic) means of the stress-point
    viscosities and of ronurestrict against the reference (g9): level-0 and
    level-1 residual rows to 1e-12 of the row scale, the restricted coarse
    viscosity bitwise."""
    g = golden("g1_operator.npz")
    r = golden("g9_viscmod.npz")
    L0, L1 = _levels(g)
    et0 = oracle.eta(L0, g["nu"], g["nxx"], g["nxy"], g["nxz"], g["nyz"], vm, float(g["nubeg"]), float(g["nuend"]))
    nu1 = oracle.ronurestrict(L0, L1, g["nu"], vm)
    assert np.array_equal(nu1, r[f"nu1_{vm}"])
    et1 = oracle.eta(L1, nu1, None, None, None, None, vm, float(g["nubeg"]), float(g["nuend"]))
    for et, L, x, b, rref, ex in ((et0, L0, r[f"x0_{vm}"], r[f"rhs0_{vm}"], r[f"r0_{vm}"], r[f"exists0_{vm}"]),
                                  (et1, L1, r[f"x1_{vm}"], r[f"b1_{vm}"], r[f"r1_{vm}"], r[f"exists1_{vm}"])):
        res = oracle.residual(L, et, g["bondm"], g["bondv"], g["bondn"], b, x)
        for c in range(4):
            d = np.abs(res[c::4] - rref[c::4])[ex[c::4]]
            s = np.abs(rref[c::4]).max() + np.abs(b[c::4]).max()
            assert d.max() <= 1e-12 * s, (vm, c, d.max(), s)
    # the means differ from the arithmetic one (a plausible slip: ignoring viscmod)
    et00 = oracle.eta(L0, g["nu"], g["nxx"], g["nxy"], g["nxz"], g["nyz"], 0, float(g["nubeg"]), float(g["nuend"]))
    assert not np.array_equal(et0[1], et00[1])


def _phase_nodes(nx, ny, nz, phase):
    """interleaved indices of the rows relaxed by one colour phase"""
    k, i, j = np.meshgrid(np.arange(nz), np.arange(nx), np.arange(ny), indexing="ij")
    n = (k * nx * ny + i * ny + j).ravel()
    if phase == 6:
        return n * 4
    sel = ((i + j + k).ravel() & 1) == (phase & 1)
    return n[sel] * 4 + 1 + phase // 2


@pytest.mark.parametrize("phase", range(7))
def test_colour_phase_matches_reference_rule(phase):
    """One phase of the colour-ordered sweep against gausmat4's update rule
    (move3mg.c:3100-3128) evaluated from the reference's own assembled rows
    (g1 residual r0 = b - A x0, g12 row coefficient a_ii and nd): velocity
    x += v0koef r / a_ii on one colour of one component, pressure
    P += p0koef nd r on every continuity row; every other entry unchanged
    bitwise."""
    g = golden("g1_operator.npz")
    s = golden("g12_smoother.npz")
    L0, _ = _levels(g)
    et = _etas(g, L0, 0, g["nu"])
    v0, p0 = float(s["v0koef"]), float(s["p0koef"])
    x0, r0 = g["x0"], g["r0"]
    out = oracle.colour_phase(L0, et, g["bondm"], g["rhs0"], x0, phase, p0, v0)
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    rows = _phase_nodes(nx, ny, nz, phase)
    bc = g["bondm"][:4 * nx * ny * nz] != 0
    rows = rows[g["exists0"][rows] & ~bc[rows]]
    if phase == 6:
        want = p0 * s["nd0"][rows // 4].astype(np.float64) * r0[rows]
    else:
        want = v0 * r0[rows] / s["diag0"][rows]
    got = out[rows] - x0[rows]
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), np.abs(got - want).max()
    others = np.ones(len(x0), bool)
    others[rows] = False
    assert np.array_equal(out[others], x0[others])
    # the reference rule differs from a sign slip or the other colour
    assert np.abs(got + want).max() > 0.5 * np.abs(want).max()


def test_lexicographic_sweep_matches_reference_gausmat4():
    """or_gs_sweep (lexicographic, the four rows of a node in P, vx, vy, vz
    order) equals one sweep of the reference's gausmat4 on level 0 (g12, one
    thread) from g1's x0."""
    g = golden("g1_operator.npz")
    s = golden("g12_smoother.npz")
    L0, _ = _levels(g)
    et = _etas(g, L0, 0, g["nu"])
    out = oracle.gs_sweep(L0, et, g["bondm"], g["bondv"], g["bondn"], g["rhs0"], g["x0"], float(s["p0koef"]),
                          float(s["v0koef"]))
    ref = s["gs_out"]
    for c in range(4):
        d = np.abs(out[c::4] - ref[c::4]).max()
        assert d <= 1e-10 * np.abs(ref[c::4] - g["x0"][c::4]).max(), (c, d)


def test_wnorms_closed_form():
    """The convergence metric's weights: a unit residual on one continuity
    row gives exactly eta_C / mean(dx, dy, dz) of its cell, on a momentum row
    1, on a BC row 0; b enters the same way."""
    g = golden("g1_operator.npz")
    L0, _ = _levels(g)
    et = _etas(g, L0, 0, g["nu"])
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    N = nx * ny * nz
    bc = g["bondm"][:4 * N] != 0
    ex = g["exists0"]
    gx, gy, gz = g["gx"], g["gy"], g["gz"]
    rng = np.random.default_rng(5)
    zero = np.zeros(4 * N)
    cont = np.where(ex & ~bc & (np.arange(4 * N) % 4 == 0))[0]
    for u in rng.choice(cont, 5, replace=False):
        n = u // 4
        k, rem = divmod(n, nx * ny)
        i, j = divmod(rem, ny)
        h = ((gx[i] - gx[i - 1]) + (gy[j] - gy[j - 1]) + (gz[k] - gz[k - 1])) / 3.0
        e = zero.copy()
        e[u] = 1.0
        w = oracle.wnorms(L0, et[0], g["bondm"], zero, e)
        assert w[0] == pytest.approx(et[0][n] / h, rel=1e-14) and w[1] == 0
        wb = oracle.wnorms(L0, et[0], g["bondm"], e, zero)
        assert wb[1] == pytest.approx(et[0][n] / h, rel=1e-14) and wb[0] == 0
    mom = np.where(ex & ~bc & (np.arange(4 * N) % 4 != 0))[0]
    e = zero.copy()
    e[mom[7]] = 1.0
    assert oracle.wnorms(L0, et[0], g["bondm"], zero, e)[0] == 1.0
    e = zero.copy()
    e[np.where(bc & ex)[0][3]] = 1.0
    assert oracle.wnorms(L0, et[0], g["bondm"], zero, e)[0] == 0.0
