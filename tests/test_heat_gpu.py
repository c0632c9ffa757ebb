"""GPU parity tests of the heat path (implicit temperature step and the
marker temperature update), through the C ABI.

The heat solve is floating point: the converged GPU temperature must match
the reference's converged titerate output (g6, written from the compiled
reference) and the oracle's sparse direct solve to 1e-7 relative L2 (the
fields are stored in float, whose rounding is ~6e-8).  The marker update
is per-marker float arithmetic on the same inputs: bit-exact against the
reference output when the node temperatures are the reference's."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

import i3elvis_amd as I
import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _heat(g):
    return I.Heat(g["gx"], g["gy"], g["gz"], g["bondm"], g["bondv"], g["bondn"])


def test_heat_step_matches_reference(gpu):
    g = golden("g6_heat.npz")
    print("case: g6 heat step", flush=True)
    H = _heat(g)
    tk, st = H.solve(g["tk"], g["ro"], g["cp"], g["kt"], g["ht"], timestep=float(g["dt"]),
                     maxtkstep=float(g["maxtkstep"]), heatmin=1e-8, sweeps=10)
    assert st.converged and st.substeps == 3
    ref = g["tk_out"]
    rel = np.linalg.norm(tk.astype(np.float64) - ref) / np.linalg.norm(ref)
    print("substeps", st.substeps, "sweeps", st.sweeps, "rel L2", rel, flush=True)
    assert rel < 1e-7
    assert np.abs(tk - ref).max() < 2e-4
    H.close()


def test_heat_default_tolerance_and_zero_step(gpu):
    g = golden("g6_heat.npz")
    H = _heat(g)
    tk, st = H.solve(g["tk"], g["ro"], g["cp"], g["kt"], g["ht"], timestep=float(g["dt"]),
                     maxtkstep=float(g["maxtkstep"]))
    assert st.converged and st.heatsum <= 1e-3
    tk0, st0 = H.solve(g["tk"], g["ro"], g["cp"], g["kt"], g["ht"], timestep=0.0)
    assert st0.substeps == 0 and np.array_equal(tk0, g["tk"])
    H.close()


def test_heat_41_matches_direct_solve(gpu):
    """41^3 non-uniform grid, random conductivity and sources, one step:
    GPU red-black relaxation against the oracle rows solved directly."""
    n = 41
    rng = np.random.default_rng(41)
    gx = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, n - 1)]) * 2.5e3
    gy = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, n - 1)]) * 2.5e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, n - 1)]) * 2.5e3
    N = n ** 3
    h = I.hierarchy(n, n, n, 0)
    rules, nr = I.freeslip_rules(273.0, 1600.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    ro = rng.uniform(2800, 3400, N).astype(np.float32)
    cp = rng.uniform(900, 1300, N).astype(np.float32)
    kt = (10 ** rng.uniform(0, 1, N)).astype(np.float32)
    ht = rng.uniform(0, 3e-6, N).astype(np.float32)
    m2 = np.tile(np.arange(n), n * n)
    tk = (273 + 1327 * gy[m2] / gy[-1] + rng.uniform(-30, 30, N)).astype(np.float32)
    dt = 3e12
    print("case: heat 41^3", flush=True)
    Hd = I.Heat(gx, gy, gz, bondm, bondv, bondn)
    out, st = Hd.solve(tk, ro, cp, kt, ht, timestep=dt, heatmin=1e-8, sweeps=20)
    assert st.converged and st.substeps == 1
    S = oracle.heat_sys(gx, gy, gz, ro, cp, kt, ht, tk, dt, bondm, bondv, bondn)
    A, b = oracle.heat_matrix(S)
    x = spl.spsolve(A.tocsc(), b)
    rel = np.linalg.norm(out - x) / np.linalg.norm(x)
    print("sweeps", st.sweeps, "rel L2", rel, flush=True)
    assert rel < 1e-7
    hs, bs = oracle.heat_norms(S, out.astype(np.float64))
    assert hs < 1e-3 and bs < 1e-4
    Hd.close()


@pytest.mark.parametrize("order", [0, 1])
def test_marker_tk_update_matches_reference(gpu, order):
    """subgrid diffusion and the nodal average of the changes: the
    reference's serial order (0) and the cell-segmented average (1, FP64
    sums reordered) against the reference's markers"""
    g = golden("g6_heat.npz")
    m = len(g["x"])
    print(f"case: marker tk update order {order}", m, flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], m, order=order)
    ids = np.arange(m, dtype=np.int32)
    perm = np.random.default_rng(6).permutation(m)
    M.upload(g["x"][perm], g["y"][perm], g["z"][perm], g["markk"][perm], np.zeros(m, np.float32),
             g["mtype"][perm], ids[perm])
    M.sort()
    M.update_tk(g["tk_out"], g["tk"], g["kt"], g["ro"], g["cp"], float(g["heatdif"]), float(g["dt"]),
                float(g["zdeep"]), float(g["tdeep"]))
    d = M.download()
    got = np.empty(m, np.float32)
    got[d["id"]] = d["tk"]
    bad = got != g["markk_out"]
    print("markers differing:", int(bad.sum()), flush=True)
    # the device exp may differ from libm by an ulp inside the diffusion factor
    assert np.allclose(got, g["markk_out"], rtol=1e-6, atol=0)
    M.close()


def test_marker_tk_update_no_diffusion_unsorted(gpu):
    """heatdif = 0 needs no sort: compare with the oracle."""
    g = golden("g6_heat.npz")
    m = len(g["x"])
    M = I.Markers(g["gx"], g["gy"], g["gz"], m)
    M.upload(g["x"], g["y"], g["z"], g["markk"], np.zeros(m, np.float32), g["mtype"])
    M.update_tk(g["tk_out"], g["tk"], g["kt"], g["ro"], g["cp"], 0.0, float(g["dt"]), float(g["zdeep"]),
                float(g["tdeep"]))
    d = M.download()
    ref = oracle.mk_tk_update(g["x"], g["y"], g["z"], g["mtype"], g["markk"], g["gx"], g["gy"], g["gz"],
                              g["tk_out"], g["tk"], g["kt"], g["ro"], g["cp"], 0.0, float(g["dt"]),
                              float(g["zdeep"]), float(g["tdeep"]))
    assert np.array_equal(d["tk"], ref)
    M.close()


def _g8_sources(g):
    return dict(gk=g["gk"], ysize=float(g["ysize"]), stp100=float(g["stp100"]), vx=g["vx"], vy=g["vy"], vz=g["vz"],
                pr=g["pr"], et=g["et"], sig=[g[k] for k in ("sxx", "syy", "szz", "sxy", "sxz", "syz")],
                eps=[g[k] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")])


def test_heat_step_with_sources_matches_reference(gpu):
    """titerate with adiabatic (adiabyn 2) and shear (frictyn 1) heating
    (g8, written from the compiled reference): same three sub-steps, the
    converged temperature within 1e-7 relative L2."""
    g = golden("g8_heat_sources.npz")
    print("case: g8 heat step with sources", flush=True)
    H = _heat(g)
    H.set_sources(2, 1, **_g8_sources(g))
    tk, st = H.solve(g["tk"], g["ro"], g["cp"], g["kt"], g["ht"], timestep=float(g["dt"]),
                     maxtkstep=float(g["maxtkstep"]), heatmin=1e-8, sweeps=10)
    ref = g["tk_out"]
    rel = np.linalg.norm(tk.astype(np.float64) - ref) / np.linalg.norm(ref)
    print("substeps", st.substeps, "sweeps", st.sweeps, "rel L2", rel, "max", np.abs(tk - ref).max(), flush=True)
    assert st.converged and st.substeps == 3
    assert rel < 1e-7
    assert np.abs(tk - ref).max() < 2e-4
    H.close()


def test_heat_sources_adiabyn1_and_off_match_oracle(gpu):
    """adiabyn 1 (alpha T rho g.v) against the oracle's source and a
    direct solve of its rows; then the terms switched off again give the
    plain step."""
    g = golden("g8_heat_sources.npz")
    H = _heat(g)
    src_kw = _g8_sources(g)
    s6, e6 = src_kw["sig"], src_kw["eps"]
    for a, fr in ((1, 0), (0, 0)):
        print(f"case: g8 sources adiabyn={a} frictyn={fr}", flush=True)
        H.set_sources(a, fr, **src_kw)
        dt = 3e12
        tk, st = H.solve(g["tk"], g["ro"], g["cp"], g["kt"], g["ht"], timestep=dt, maxtkstep=1e9, heatmin=1e-8,
                         sweeps=10)
        src, rate = oracle.heat_source(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["ht"], g["tk"], g["et"], g["vx"],
                                       g["vy"], g["vz"], g["pr"], s6, e6, g["bondm"], a, fr, gk=g["gk"],
                                       ysize=float(g["ysize"]), stp100=float(g["stp100"]))
        Hs = oracle.heat_sys(g["gx"], g["gy"], g["gz"], g["ro"], g["cp"], g["kt"], g["ht"], g["tk"], dt, g["bondm"],
                             g["bondv"], g["bondn"], src=src)
        A, b = oracle.heat_matrix(Hs)
        ref = spl.spsolve(A.tocsc(), b)
        rel = np.linalg.norm(tk - ref) / np.linalg.norm(ref)
        print("  substeps", st.substeps, "rel L2", rel, flush=True)
        assert st.converged and st.substeps == 1
        assert rel < 1e-7
    H.close()
