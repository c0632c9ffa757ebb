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
