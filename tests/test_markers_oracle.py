"""CPU tests: marker oracle functions against the reference's own outputs
(g5 fixture written from the compiled reference) and closed forms."""
import numpy as np

import oracle
from conftest import golden


def _rocks(g):
    return [dict(zip(oracle.ROCK_FIELDS, r)) for r in g["rocks"]]


def test_cell_search_bitexact():
    g = golden("g5_markers.npz")
    p = g["pts"]
    c = oracle.mk_cells(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), g["gx"], g["gy"], g["gz"])
    assert np.array_equal(c, g["cells"])


def test_cell_search_closed_form():
    """largest i with g[i] <= clamp(x), clamped to n-2 (searchsorted form)"""
    rng = np.random.default_rng(3)
    g = np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 40)])
    x = np.r_[rng.uniform(-2, g[-1] + 2, 5000), g].astype(np.float32)
    c = oracle.mk_cells(x, x, x, g, g, g)[:, 0]
    xc = np.clip(x.astype(np.float64), 0, g[-1])
    ref = np.clip(np.searchsorted(g, xc, side="right") - 1, 0, len(g) - 2)
    assert np.array_equal(c, ref)


def test_ronurecalc_matches_reference():
    g = golden("g5_markers.npz")
    fld, eta = oracle.ronurecalc(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["gx"], g["gy"], g["gz"],
                                 g["pr"], _rocks(g), g["init"], float(g["nubeg"]), float(g["nuend"]), 0.0, 1,
                                 float(g["timestep"]), float(g["vdeep"]), float(g["nudeep"]))
    fld[6] = oracle.tkrecalc(fld[6], g["bondm"], g["bondv"], g["bondn"])
    assert np.array_equal(eta, g["eta_out"])
    for q in range(11):
        assert np.array_equal(fld[q], g["out"][q]), q


def test_vxyz_matches_reference():
    g = golden("g5_markers.npz")
    for p, vref in zip(g["pts"], g["vout"]):
        v = oracle.vxyz(g["vx"], g["vy"], g["vz"], g["gx"], g["gy"], g["gz"], float(p[0]), float(p[1]), float(p[2]))
        assert np.array_equal(v, vref)


def test_rk4_uniform_flow_translates():
    """constant velocity: every RK4 stage is the same, x += dt*v exactly in
    double before the float store"""
    gx = np.linspace(0, 1e5, 12)
    N = 12 ** 3
    vx = np.full(N, 1e-9, np.float32)
    vy = np.full(N, -2e-9, np.float32)
    vz = np.full(N, 0.5e-9, np.float32)
    rng = np.random.default_rng(5)
    x, y, z = (rng.uniform(2e4, 8e4, 100).astype(np.float32) for _ in range(3))
    dt = 1e12
    out = oracle.rk4(x, y, z, np.ones(100, np.float32), np.ones(100, np.float32), np.full(100, 3, np.int8),
                     vx, vy, vz, gx, gx, gx, dt)
    assert np.array_equal(out[0], (x.astype(np.float64) + dt * np.float64(np.float32(1e-9))).astype(np.float32))
    assert np.array_equal(out[1], (y.astype(np.float64) + dt * np.float64(np.float32(-2e-9))).astype(np.float32))


def test_epscalc_matches_reference_bitwise():
    """epscalc1 (strain rates, stresses, P, dissipation at a point with
    clamped trilinear weights) against the reference at 600 points of the
    g5 grid, including points outside the grid (clamped) and on grid and
    mid lines (g10): equal to the bit."""
    g = golden("g5_markers.npz")
    r = golden("g10_eps_movemark.npz")
    st = [r[k] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")]
    out = np.array([oracle.epscalc(st, r["pr"], g["gx"], g["gy"], g["gz"], *p) for p in r["pts"]])
    assert np.array_equal(out, r["eps"])


def test_epscalc_closed_forms():
    """Constant fields are reproduced exactly inside the grid (weights of
    each block sum to one), and the dissipation of constant fields is
    sigma:eps with the shear terms doubled."""
    g = golden("g5_markers.npz")
    N = len(g["pr"])
    vals = [1e-15, 2e-15, -3e-15, 4e-15, 5e-15, -6e-15, 1e6, 2e6, 3e6, -4e6, 5e6, 6e6]
    st = [np.full(N, v, np.float32) for v in vals]
    pr = np.full(N, 3e8)
    rng = np.random.default_rng(2)
    for _ in range(50):
        p = [rng.uniform(g[k][2], g[k][-3]) for k in ("gx", "gy", "gz")]
        out = oracle.epscalc(st, pr, g["gx"], g["gy"], g["gz"], *p)
        f = np.float32(vals)
        assert np.allclose(out[:6], f[:6], rtol=1e-12)
        assert np.isclose(out[6], 3e8, rtol=1e-12)
        assert np.allclose(out[7:13], f[6:], rtol=1e-12)
        dis = float(f[6] * f[0] + f[7] * f[1] + f[8] * f[2]) + 2 * sum(float(f[9 + q]) * float(f[3 + q]) for q in range(3))
        assert np.isclose(out[13], dis, rtol=1e-6)


def test_movemark_with_depth_rules_matches_reference():
    """movemark of the g5 markers with zdeep/vdeep inside the grid (g11,
    written from the reference): the oracle's property update and RK4 with
    the depth rules reproduce viscosity, strain, positions, temperatures
    and the changed types bit for bit."""
    g = golden("g5_markers.npz")
    r10 = golden("g10_eps_movemark.npz")
    r = golden("g11_movemark_depth.npz")
    st = [r10[k] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")]
    rocks = [dict(zip(oracle.ROCK_FIELDS, row)) for row in g["rocks"]]
    dt = float(r["dt"])
    eta, e = oracle.mk_props(g["x"], g["y"], g["z"], g["mtype"], g["tk"], r10["markv"], r10["marke"], rocks, st, r10["pr"],
                             g["gx"], g["gy"], g["gz"], float(g["nubeg"]), float(g["nuend"]), dt, float(r["healrate"]))
    x, y, z, tk, eta, t = oracle.rk4(g["x"], g["y"], g["z"], g["tk"], eta, g["mtype"], r["vx"], r["vy"], r["vz"], g["gx"],
                                     g["gy"], g["gz"], dt, depth=(float(r["zdeep"]), float(r["vdeep"]), float(r["tdeep"])))
    for a, b in ((x, r["x_out"]), (y, r["y_out"]), (z, r["z_out"]), (tk, r["tk_out"]), (eta, r["eta_out"]),
                 (e, r["e_out"]), (t, r["type_out"])):
        assert np.array_equal(a, b)
    assert (t != g["mtype"]).sum() > 100
