"""CPU tests: heat oracle (rows, step limit, norms, marker temperature
update) against the reference's own outputs (g6 fixture written from the
compiled reference) and closed forms."""
import numpy as np
import scipy.sparse.linalg as spl

import oracle
from conftest import golden


def heat_substeps(g, dt, maxtkstep):
    """titerate's sub-step sequence (heat3mg.c:358-384)."""
    H = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], g["tk"], dt,
                        g["bondm"], g["bondv"], g["bondn"])
    rate = oracle.heat_maxrate(H)
    step = maxtkstep / rate if rate else 1e30
    out, s = [], 0.0
    while s < dt:
        h = min(dt, step)
        if h > dt - s:
            h = dt - s
        out.append(h)
        s += h
    return out


def oracle_heat(g):
    """Sub-stepped implicit heat solve by a sparse direct solve of the
    oracle's rows; T stored as float after every sub-step."""
    tk = g["tk"].astype(np.float32)
    for h in heat_substeps(g, float(g["dt"]), float(g["maxtkstep"])):
        H = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], tk, h,
                            g["bondm"], g["bondv"], g["bondn"])
        A, b = oracle.heat_matrix(H)
        tk = spl.spsolve(A.tocsc(), b).astype(np.float32)
    return tk


def test_heat_substeps_three():
    g = golden("g6_heat.npz")
    st = heat_substeps(g, float(g["dt"]), float(g["maxtkstep"]))
    assert len(st) == 3 and np.isclose(sum(st), float(g["dt"]), rtol=1e-15)


def test_heat_solve_matches_reference():
    g = golden("g6_heat.npz")
    tk = oracle_heat(g)
    ref = g["tk_out"]
    assert np.linalg.norm(tk.astype(np.float64) - ref) / np.linalg.norm(ref) < 1e-7
    assert np.abs(tk - ref).max() < 2e-4


def test_heat_norms_of_converged_solution_small():
    g = golden("g6_heat.npz")
    H = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], g["tk"], 1e12,
                        g["bondm"], g["bondv"], g["bondn"])
    A, b = oracle.heat_matrix(H)
    x = spl.spsolve(A.tocsc(), b)
    hs, bs = oracle.heat_norms(H, x)
    assert bs < 1e-9
    assert hs < 1e-3          # float storage floor of the reference norm
    hs0, _ = oracle.heat_norms(H, g["tk"].astype(np.float64))
    assert hs0 > 100 * hs


def test_heat_linear_profile_is_steady():
    """constant k, no sources, T linear in depth between fixed top/bottom and
    zero-flux sides: the implicit step leaves T unchanged."""
    n = 9
    g = np.linspace(0, 8e3, n)
    N = n ** 3
    import i3elvis_amd as I
    h = I.hierarchy(n, n, n, 0)
    rules, nr = I.freeslip_rules(300.0, 1100.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    m2 = np.tile(np.arange(n), n * n)
    tk = (300.0 + 800.0 * m2 / (n - 1)).astype(np.float32)
    one = np.ones(N, np.float32)
    H = oracle.heat_sys(g, g, g, 3300 * one, 1000 * one, 3 * one, 0 * one, tk, 1e13, bondm, bondv, bondn)
    A, b = oracle.heat_matrix(H)
    x = spl.spsolve(A.tocsc(), b)
    assert np.abs(x - tk).max() < 1e-9


def test_heat_fourier_mode_decay():
    """single cosine mode in x with zero-flux sides in x and z and fixed T=0
    top/bottom replaced by a y-uniform mode: one implicit Euler step scales
    the discrete mode by 1/(1 + dt*kappa*lambda_h)."""
    n = 17
    L = 16e3
    g = np.linspace(0, L, n)
    import i3elvis_amd as I
    N = n ** 3
    h = I.hierarchy(n, n, n, 0)
    rules, nr = I.freeslip_rules(0.0, 0.0)
    rules = [r for r in rules[:nr] if r.var == 4 and not (r.m20 == 0 and r.m21 == 0 and r.rel20 == 0)
             and not (r.rel20 == 1 and r.rel21 == 1)]
    # keep only the side rules (x and z walls); y walls become zero-flux too
    import ctypes as C
    extra = (I.BcRule * 2)()
    for k, (lo, rel, d) in enumerate(((0, 0, 1), (0, 1, -1))):
        r = extra[k]
        r.var = 4; r.m10, r.m11, r.rel10, r.rel11 = 0, 0, 0, 1
        r.m20 = r.m21 = lo; r.rel20 = r.rel21 = rel
        r.m30, r.m31, r.rel30, r.rel31 = 0, 0, 0, 1
        r.cnst, r.koef, r.dm1, r.dm2, r.dm3 = 0.0, 1.0, 0, d, 0
    allr = (I.BcRule * (len(rules) + 2))(*rules, *extra)
    bondm, bondv, bondn = I.expand_bc(allr, len(allr), h)
    m1 = (np.arange(N) // n) % n
    dx = L / (n - 1)
    # ghost-free cosine on interior nodes 1..n-2 with mirrored walls: cos(pi*(m1-0.5)/(n-2))
    k = np.pi / ((n - 2) * dx)
    T0 = np.cos(k * (m1 - 0.5) * dx)
    kap = 3.0 / (3300 * 1000)
    dt = 1e12
    one = np.ones(N, np.float32)
    H = oracle.heat_sys(g, g, g, 3300 * one, 1000 * one, 3 * one, 0 * one, T0.astype(np.float32), dt,
                        bondm, bondv, bondn)
    A, b = oracle.heat_matrix(H)
    x = spl.spsolve(A.tocsc(), b)
    lam = (2 - 2 * np.cos(k * dx)) / dx ** 2
    expect = T0.astype(np.float32).astype(np.float64) / (1 + dt * kap * lam)
    inner = (m1 > 0) & (m1 < n - 1)
    assert np.abs(x[inner] - expect[inner]).max() < 1e-6


def test_marker_tk_update_matches_reference():
    g = golden("g6_heat.npz")
    tk = g["tk_out"]
    mk = oracle.mk_tk_update(g["x"], g["y"], g["z"], g["mtype"], g["markk"], g["gx"], g["gy"], g["gz"], tk,
                             g["tk"], g["kt"], g["ro"], g["cp"], float(g["heatdif"]), float(g["dt"]),
                             float(g["zdeep"]), float(g["tdeep"]))
    assert np.array_equal(mk, g["markk_out"])


def test_marker_tk_update_no_diffusion_is_interpolated_change():
    """heatdif = 0: every marker inside the grid gains the trilinear
    interpolation of T_new - T_start, which for a linear field is exact."""
    g = golden("g6_heat.npz")
    m1 = np.arange(len(g["tk"])) // len(g["gy"]) % len(g["gx"])
    dT = (5.0 + 0.25 * m1).astype(np.float32)
    tk2 = g["tk"].astype(np.float32)
    tk = tk2 + dT
    mk0 = np.abs(g["markk"]) + 1.0
    mk = oracle.mk_tk_update(g["x"], g["y"], g["z"], np.full(len(mk0), 3, np.int8), mk0, g["gx"], g["gy"], g["gz"],
                             tk, tk2, g["kt"], g["ro"], g["cp"], 0.0, 1.0)
    inside = g["x"] > 0
    gx = g["gx"]
    xi = np.interp(g["x"][inside], gx, np.arange(len(gx)))
    expect = mk0[inside] + (5.0 + 0.25 * xi)
    assert np.abs(mk[inside] - expect).max() < 1e-3


# ---- adiabatic and shear heating (g8) --------------------------------------

def _g8_source(g, a, fr, **kw):
    s6 = [g[k] for k in ("sxx", "syy", "szz", "sxy", "sxz", "syz")]
    e6 = [g[k] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")]
    args = dict(gk=g["gk"], ysize=float(g["ysize"]), stp100=float(g["stp100"]))
    args.update(kw)
    return oracle.heat_source(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["ht"], g["tk"], g["et"], g["vx"], g["vy"],
                              g["vz"], g["pr"], s6, e6, g["bondm"], a, fr, **args)


def test_heat_source_matches_reference_bitwise():
    """heaterr ynerr=0 of the compiled reference (g8): the stored tk1 and
    the returned rate at every interior node, for adiabyn 1 and 2 and
    frictyn 1, equal to the bit."""
    g = golden("g8_heat_sources.npz")
    for a, fr in ((1, 0), (2, 1), (0, 1)):
        src, rate = _g8_source(g, a, fr)
        assert np.array_equal(src, g[f"src_{a}{fr}"]), (a, fr)
        assert np.array_equal(rate, g[f"rate_{a}{fr}"]), (a, fr)


def test_heat_source_terms_switch_and_mask():
    """Without adiabyn/frictyn the source is the radiogenic heat; the terms
    act only where rho > 2000 and y < ysize - stp100 (plausible slips --
    a dropped mask or a wrong sign of the gradient -- change these)."""
    g = golden("g8_heat_sources.npz")
    N = len(g["ro"])
    bt = g["bondm"][8 * N:9 * N]
    inner = bt == 0
    src0, _ = _g8_source(g, 0, 0)
    assert np.array_equal(src0[inner], g["ht"][inner])
    n = int(g["n"])
    y = np.tile(g["gy"], n * n)
    off = inner & ((g["ro"] <= 2000) | (y >= float(g["ysize"]) - float(g["stp100"])))
    assert off.sum() > 100 and (inner & ~off).sum() > 100
    for a, fr in ((1, 0), (2, 1)):
        src, _ = _g8_source(g, a, fr)
        assert np.array_equal(src[off], g["ht"][off])
        assert np.abs(src[inner & ~off] - g["ht"][inner & ~off]).max() > 0


def test_adiabatic_term_closed_forms():
    """adiabyn 1 in a uniform upward flow: alpha T rho g.v exactly; adiabyn
    2 with a linear pressure P = c y and the same flow: alpha T c v_y."""
    g = golden("g8_heat_sources.npz")
    N = len(g["ro"])
    bt = g["bondm"][8 * N:9 * N]
    n = int(g["n"])
    y = np.tile(g["gy"], n * n)
    one = np.ones(N, np.float32)
    sel = (bt == 0) & (y < float(g["ysize"]) - float(g["stp100"]))
    z6 = [0 * one] * 6
    vy = 2e-9
    ht = 0 * one
    src1, _ = oracle.heat_source(g["gx"], g["gy"], g["gz"], 3000 * one, 1000 * one, ht, 1000 * one, 3e-5 * one, 0 * one,
                                 vy * one, 0 * one, 0 * y, z6, z6, g["bondm"], 1, 0, gk=(0.0, 9.81, 0.0),
                                 ysize=float(g["ysize"]), stp100=float(g["stp100"]))
    expect1 = np.float32(np.float32(3e-5) * np.float32(1000) * np.float32(3000)) * (vy * 9.81)
    assert np.allclose(src1[sel], expect1, rtol=1e-6)
    c = 3.2e4
    src2, _ = oracle.heat_source(g["gx"], g["gy"], g["gz"], 3000 * one, 1000 * one, ht, 1000 * one, 3e-5 * one, 0 * one,
                                 vy * one, 0 * one, c * y, z6, z6, g["bondm"], 2, 0, gk=(0.0, 9.81, 0.0),
                                 ysize=float(g["ysize"]), stp100=float(g["stp100"]))
    # the P nodes sit at cell centres: (y_c(m2+1) - y_c(m2)) ... c y is linear, so dP/dy = c in the interior
    m2 = np.tile(np.arange(n), n * n)
    sel2 = sel & (m2 < n - 2)
    assert np.allclose(src2[sel2], np.float32(3e-5 * 1000.0) * c * vy, rtol=1e-6)


def test_heat_solve_with_sources_matches_reference():
    """titerate with adiabyn 2 and frictyn 1 (g8): the oracle rows with the
    oracle source (tk1) and the sub-steps from its rate reproduce the
    reference's temperature."""
    g = golden("g8_heat_sources.npz")
    src, rate = _g8_source(g, 2, 1)
    dt, maxtkstep = float(g["dt"]), float(g["maxtkstep"])
    step = maxtkstep / np.abs(rate).max()
    subs, s = [], 0.0
    while s < dt:
        h = min(dt, step)
        h = min(h, dt - s)
        subs.append(h)
        s += h
    assert len(subs) == 3
    tk = g["tk"].astype(np.float32)
    for h in subs:
        H = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], tk, h, g["bondm"], g["bondv"],
                            g["bondn"], src=src)
        A, b = oracle.heat_matrix(H)
        tk = spl.spsolve(A.tocsc(), b).astype(np.float32)
    ref = g["tk_out"]
    assert np.linalg.norm(tk.astype(np.float64) - ref) / np.linalg.norm(ref) < 1e-7
    assert np.abs(tk - ref).max() < 2e-4
    # the sources matter: without them the step differs clearly
    tk0 = g["tk"].astype(np.float32)
    for h in subs:
        H = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], tk0, h, g["bondm"],
                            g["bondv"], g["bondn"])
        A, b = oracle.heat_matrix(H)
        tk0 = spl.spsolve(A.tocsc(), b).astype(np.float32)
    assert np.abs(tk0 - ref).max() > 100 * np.abs(tk - ref).max()
