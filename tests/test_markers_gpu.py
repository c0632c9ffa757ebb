"""GPU parity tests of the marker path, through the C ABI.

Cell indices and the sort order must be bit-exact (integer work).  The
marker -> node fields in the reference's serial accumulation order (order 0)
are expected to agree to the last float bit except where the device's exp
differs from libm by an ulp (Arrhenius viscosities): tolerance 2e-7
relative, with the bit-exact fraction reported.  The cell-segmented
transfer (order 1, the default) sums in FP64: against the oracle's FP64
node sums 2e-7; against the reference's float accumulation within that
accumulation's own rounding, k 2^-24 for k weighted terms per node (4e-6
here), and for the geometric mean (sums of ln(eta)) that bound times the
largest |ln eta|.
"""
import numpy as np
import pytest

import i3elvis_amd as I
import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


def _markers(g, order=0):
    M = I.Markers(g["gx"], g["gy"], g["gz"], len(g["x"]) + 16, order=order)
    M.upload(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"])
    return M


def test_cell_index_bitexact(gpu):
    g = golden("g5_markers.npz")
    print("case: g5 cell index", len(g["x"]), flush=True)
    M = _markers(g)
    m1, m2, m3, key = M.cell_index()
    c = oracle.mk_cells(g["x"], g["y"], g["z"], g["gx"], g["gy"], g["gz"])
    assert np.array_equal(np.stack([m1, m2, m3], 1), c)
    nx, ny = len(g["gx"]), len(g["gy"])
    assert np.array_equal(key, c[:, 2] * nx * ny + c[:, 0] * ny + c[:, 1])
    M.close()


def test_sort_is_stable_by_cell_then_id(gpu):
    g = golden("g5_markers.npz")
    rng = np.random.default_rng(9)
    perm = rng.permutation(len(g["x"]))
    M = I.Markers(g["gx"], g["gy"], g["gz"], len(g["x"]) + 16)
    ids = np.arange(len(g["x"]), dtype=np.int32)
    M.upload(g["x"][perm], g["y"][perm], g["z"][perm], g["tk"][perm], g["eta"][perm], g["mtype"][perm], ids[perm])
    key0 = M.cell_index()[3]
    order = M.sort()
    d = M.download()
    expect = np.lexsort((ids[perm], key0))
    assert np.array_equal(order, expect)
    assert np.array_equal(d["id"], ids[perm][expect])
    assert np.array_equal(d["x"], g["x"][perm][expect])
    # sorting again is the identity
    assert np.array_equal(M.sort(), np.arange(len(order)))
    M.close()


def test_sort_cells_of_every_size(gpu):
    """The in-cell ordering by id for cells of 1 to 64 markers (bitonic
    network across the lanes), 65 to 1024 (in LDS) and more (sorted chunks
    of 1024 merged pairwise: 2, 3, 5 and 9 chunks, so that some passes have
    an odd run left over), with ids in random order: the order equals
    numpy's lexsort by (cell, id)."""
    rng = np.random.default_rng(21)
    n = 9
    g = np.linspace(0.0, 8.0, n)
    # cell (a, b, c) receives 0..9000 markers: a few crowded cells
    sizes = {0: 1, 1: 7, 2: 64, 3: 65, 4: 200, 5: 1024, 6: 1025, 7: 1600, 8: 2048, 9: 3100, 10: 5000,
             11: 9000, 100: 33, 300: 129, 511: 4097}
    rest = rng.integers(0, 40, 512)
    pts = []
    for cell in range(512):
        k = sizes.get(cell, int(rest[cell]))
        c3, r = divmod(cell, 64)
        c1, c2 = divmod(r, 8)
        pts.append(np.stack([c1 + rng.uniform(0.01, 0.99, k), c2 + rng.uniform(0.01, 0.99, k),
                             c3 + rng.uniform(0.01, 0.99, k)], 1))
    pts = np.concatenate(pts).astype(np.float32)
    m = len(pts)
    perm = rng.permutation(m)
    ids = rng.permutation(m).astype(np.int32)
    print(f"case: sort {m} markers, cells up to 9000", flush=True)
    M = I.Markers(g, g, g, m)
    z = np.zeros(m, np.float32)
    M.upload(pts[perm, 0], pts[perm, 1], pts[perm, 2], z, z, np.full(m, 2, np.int8), ids[perm])
    key0 = M.cell_index()[3]
    order = M.sort()
    expect = np.lexsort((ids[perm], key0))
    assert np.array_equal(order, expect)
    assert np.array_equal(M.download()["id"], ids[perm][expect])
    assert np.array_equal(M.sort(), np.arange(m))
    M.close()


@pytest.mark.parametrize("order", [0, 1])
def test_to_nodes_matches_reference(gpu, order):
    g = golden("g5_markers.npz")
    print(f"case: g5 to_nodes order {order}", flush=True)
    M = _markers(g, order)
    M.sort()
    M.set_tbc(g["bondm"], g["bondv"], g["bondn"])
    rocks = [list(r) for r in g["rocks"]]
    fld = M.to_nodes(rocks, g["pr"], g["init"], float(g["nubeg"]), float(g["nuend"]), 0.0, 1,
                     float(g["timestep"]), float(g["vdeep"]), float(g["nudeep"]))
    ref = g["out"]
    exact = 0
    for q in range(11):
        a, b = fld[q], ref[q]
        rel = np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), 1e-30)
        exact += int(np.sum(a == b))
        assert rel.max() <= (2e-7 if order == 0 else 4e-6), (q, rel.max())
    print("bit-exact node values:", exact, "of", ref.size, flush=True)
    if order == 1:
        # the oracle's FP64 node sums of the same markers (tkrecalc applied
        # by the library is left out: BC nodes are compared with the golden)
        oracle.set_sum64(True)
        try:
            o, _ = oracle.ronurecalc(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["gx"], g["gy"],
                                     g["gz"], g["pr"], [dict(zip(oracle.ROCK_FIELDS, r)) for r in g["rocks"]],
                                     g["init"], float(g["nubeg"]), float(g["nuend"]), 0.0, 1, float(g["timestep"]),
                                     float(g["vdeep"]), float(g["nudeep"]))
        finally:
            oracle.set_sum64(False)
        o[6] = oracle.tkrecalc(o[6], g["bondm"], g["bondv"], g["bondn"])
        for q in range(11):
            rel = np.abs(fld[q].astype(np.float64) - o[q]) / np.maximum(np.abs(o[q]), 1e-30)
            assert rel.max() <= 2e-7, (q, rel.max())
    d = M.download()
    eta_ref = g["eta_out"][d["id"]]
    assert np.allclose(d["eta"], eta_ref, rtol=2e-7, atol=0)
    M.close()


def test_rk4_matches_oracle(gpu):
    g = golden("g5_markers.npz")
    rng = np.random.default_rng(11)
    N = len(g["pr"])
    vx, vy, vz = (rng.standard_normal(N).astype(np.float32) * 1e-9 for _ in range(3))
    M = _markers(g)
    dt = 3e11
    M.advect_rk4(dt, vx, vy, vz)
    d = M.download()
    ox, oy, oz, otk, oeta = oracle.rk4(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], vx, vy, vz,
                                       g["gx"], g["gy"], g["gz"], dt)
    for a, b in ((d["x"], ox), (d["y"], oy), (d["z"], oz)):
        ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, ulp.max()
    assert np.array_equal(d["tk"] == 0, otk == 0)
    M.close()


def _jittered(n, reps, seed):
    cfg = I.make_config(1, n, n, n, seed=3)
    gx, gy, gz = cfg["gx"], cfg["gy"], cfg["gz"]
    rng = np.random.default_rng(seed)
    cells = np.stack(np.meshgrid(np.arange(n - 1), np.arange(n - 1), np.arange(n - 1), indexing="ij"), -1).reshape(-1, 3)
    c = np.repeat(cells, reps, 0)
    x = (gx[c[:, 0]] + rng.uniform(0.01, 0.99, len(c)) * (gx[1] - gx[0])).astype(np.float32)
    y = (gy[c[:, 1]] + rng.uniform(0.01, 0.99, len(c)) * (gy[1] - gy[0])).astype(np.float32)
    z = (gz[c[:, 2]] + rng.uniform(0.01, 0.99, len(c)) * (gz[1] - gz[0])).astype(np.float32)
    perm = rng.permutation(len(x))
    mtype = rng.integers(2, 7, len(x)).astype(np.int8)
    tk = rng.uniform(600, 1700, len(x)).astype(np.float32)
    return gx, gy, gz, x[perm], y[perm], z[perm], tk, mtype


@pytest.mark.parametrize("order", [0, 1])
def test_to_nodes_matches_oracle_65(gpu, order):
    """65^3 grid, 8 shuffled markers per cell, Arrhenius rocks: GPU transfer
    against the oracle's serial restatement (float node sums for order 0,
    FP64 node sums for the cell-segmented order 1)."""
    n = 65
    gx, gy, gz, x, y, z, tk, mtype = _jittered(n, 8, 5)
    m = len(x)
    print(f"case: to_nodes 65^3 order {order}", m, flush=True)
    g = golden("g5_markers.npz")
    rocks = [list(r) for r in g["rocks"]]
    eta = np.zeros(m, np.float32)
    pr = np.random.default_rng(2).uniform(1e7, 1e9, n ** 3)
    init = np.zeros((11, n ** 3), np.float32)
    M = I.Markers(gx, gy, gz, m, order=order)
    M.upload(x, y, z, tk, eta, mtype)
    M.sort()
    fld = M.to_nodes(rocks, pr, init, 1e18, 1e25, 0.0, 1, 1e10, gy[-1] * 0.8, 1e21)
    d = M.download()
    oracle.set_sum64(order == 1)
    try:
        ref, eref = oracle.ronurecalc(x, y, z, tk, eta, mtype, gx, gy, gz, pr,
                                      [dict(zip(oracle.ROCK_FIELDS, r)) for r in rocks],
                                      init, 1e18, 1e25, 0.0, 1, 1e10, gy[-1] * 0.8, 1e21)
    finally:
        oracle.set_sum64(False)
    for q in range(11):
        rel = np.abs(fld[q].astype(np.float64) - ref[q]) / np.maximum(np.abs(ref[q]), 1e-30)
        assert rel.max() <= 2e-7, (q, rel.max())
    assert np.allclose(d["eta"], eref[d["id"]], rtol=2e-7, atol=0)
    M.close()


def test_fused_transfer_equals_staged_records_33(gpu, monkeypatch):
    """33^3 grid, 12 shuffled markers per cell (one full and one half chunk
    of eight per cell), Arrhenius rocks, viscosities computed by the
    transfer (eta 0): the one-pass transfer equals the two-pass one that
    stages the records in memory (I3G_MK_STAGED=1) to the bit."""
    n = 33
    gx, gy, gz, x, y, z, tk, mtype = _jittered(n, 12, 7)
    m = len(x)
    print(f"case: fused vs staged transfer 33^3, {m} markers", flush=True)
    rocks = [list(r) for r in golden("g5_markers.npz")["rocks"]]
    pr = np.random.default_rng(4).uniform(1e7, 1e9, n ** 3)
    init = np.zeros((11, n ** 3), np.float32)
    out = []
    for staged in ("0", "1"):
        monkeypatch.setenv("I3G_MK_STAGED", staged)
        M = I.Markers(gx, gy, gz, m)
        M.upload(x, y, z, tk, np.zeros(m, np.float32), mtype)
        M.sort()
        f = M.to_nodes(rocks, pr, init, 1e18, 1e25, 0.0, 1, 1e10, gy[-1] * 0.8, 1e21)
        out.append((f, M.download()))
        M.close()
    assert np.array_equal(out[0][0], out[1][0])
    for k in ("eta", "type", "id"):
        assert np.array_equal(out[0][1][k], out[1][1][k]), k
    assert np.isfinite(out[0][0]).all()


def test_large_marker_set_deterministic(gpu):
    """129^3 grid, 8 markers per cell (16.8M): two transfers give identical
    node fields; constant density maps to a constant."""
    n = 129
    gx, gy, gz, x, y, z, tk, mtype = _jittered(n, 8, 1)
    m = len(x)
    print("case: large marker set", m, flush=True)
    M = I.Markers(gx, gy, gz, m)
    M.upload(x, y, z, np.full(m, 1000, np.float32), np.full(m, 1e21, np.float32), np.full(m, 2, np.int8))
    M.sort()
    rocks = [[1e21, 0, 0, 1e18, 1e25, 3300, 0, 0, 1000, 3, 0, 0, 0]] * 3
    init = np.zeros((11, n ** 3), np.float32)
    pr = np.zeros(n ** 3)
    f1 = M.to_nodes(rocks, pr, init, 1e18, 1e25, densimod=0, timestep=1.0)
    f2 = M.to_nodes(rocks, pr, init, 1e18, 1e25, densimod=0, timestep=1.0)
    assert np.array_equal(f1, f2)
    # FP64 sums of ~64 weighted terms per node (cell-segmented default)
    assert np.allclose(f1[1], 3300, rtol=1e-7)
    assert np.allclose(f1[3], 3, rtol=1e-7)
    # the reference's serial order: float sums, within their rounding
    M.set_order(0)
    f0 = M.to_nodes(rocks, pr, init, 1e18, 1e25, densimod=0, timestep=1.0)
    assert np.array_equal(f0, M.to_nodes(rocks, pr, init, 1e18, 1e25, densimod=0, timestep=1.0))
    assert np.allclose(f0[1], 3300, rtol=1e-5)
    assert np.allclose(f0, f1, rtol=64 * 2.0 ** -24, atol=0)
    M.close()


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("vm", [1, 2])
def test_to_nodes_viscmod_matches_reference(gpu, vm, order):
    """ronurecalc with geometric (1) / harmonic (2) averaging of the marker
    viscosities (mark3mg.c:848-981) against the reference's node fields
    (g9, the g5 markers): within 2 float ulps-level relative error in the
    serial order; the cell-segmented order within the float accumulation's
    rounding (for the geometric mean scaled by ln eta ~ 55)."""
    g = golden("g5_markers.npz")
    r = golden("g9_viscmod.npz")
    print(f"case: g9 to_nodes viscmod={vm} order {order}", flush=True)
    M = _markers(g, order)
    M.sort()
    M.set_tbc(g["bondm"], g["bondv"], g["bondn"])
    rocks = [list(x) for x in g["rocks"]]
    fld = M.to_nodes(rocks, g["pr"], g["init"], float(g["nubeg"]), float(g["nuend"]), 0.0, 1,
                     float(g["timestep"]), float(g["vdeep"]), float(g["nudeep"]), viscmod=vm)
    ref = r[f"mk_out_{vm}"]
    tol = 4e-7 if order == 0 else (4e-6 if vm == 2 else 4e-6 * 60)
    for q in range(11):
        a, b = fld[q], ref[q]
        rel = np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), 1e-30)
        print(f"  field {q}: max rel {rel.max():.2e}", flush=True)
        assert rel.max() <= tol, (vm, q, rel.max())
    assert not np.array_equal(ref[0], g["out"][0])
    M.close()


_ST12 = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def test_epscalc_matches_oracle_bitwise(gpu):
    """i3g_markers_epscalc at 600 markers (points of g10: inside, outside
    -- clamped -- and on grid and mid lines) against the oracle's epscalc1
    restatement, itself equal to the reference to the bit: equal to the
    bit (the device uses the same summation order, nothing contracted)."""
    g = golden("g5_markers.npz")
    r = golden("g10_eps_movemark.npz")
    pts = r["pts"].astype(np.float32)
    n = len(pts)
    print("case: epscalc at", n, "markers", flush=True)
    M = I.Markers(g["gx"], g["gy"], g["gz"], n)
    M.upload(pts[:, 0], pts[:, 1], pts[:, 2], np.ones(n, np.float32), np.zeros(n, np.float32), np.full(n, 2, np.int8))
    st = [r[k] for k in _ST12]
    out = M.epscalc(st, r["pr"])
    ref = np.array([oracle.epscalc(st, r["pr"], g["gx"], g["gy"], g["gz"], *p.astype(np.float64)) for p in pts])
    d = np.abs(out - ref).max(axis=0)
    print("max |diff| per output", d, flush=True)
    assert np.array_equal(out, ref)
    M.close()


def test_movemark_props_and_rk4_match_reference(gpu):
    """One movemark step of the g5 markers through the device
    (i3g_markers_update_props, then i3g_markers_advect_rk4) against the
    reference's own movemark (g10): viscosity from viscalc at the marker
    pressure of epscalc1 (2e-7, the device exp), accumulated strain
    (bitwise), positions (1 ulp) and the out-of-grid reset."""
    g = golden("g5_markers.npz")
    r = golden("g10_eps_movemark.npz")
    print("case: movemark props + rk4", len(g["x"]), flush=True)
    M = _markers(g)
    n = len(g["x"])
    M.upload(g["x"], g["y"], g["z"], g["tk"], r["markv"], g["mtype"])
    M.set_strain(r["marke"])
    rocks = [list(x) for x in g["rocks"]]
    dt = float(r["dt"])
    M.update_props(rocks, float(g["nubeg"]), float(g["nuend"]), dt, float(r["healrate"]), [r[k] for k in _ST12], r["pr"])
    M.advect_rk4(dt, r["vx"], r["vy"], r["vz"])
    d = M.download()
    e = M.get_strain()
    assert np.array_equal(d["id"], np.arange(n))
    assert np.array_equal(e, r["e_out"])
    rel = np.abs(d["eta"].astype(np.float64) - r["eta_out"]) / np.maximum(np.abs(r["eta_out"]), 1e-30)
    print("eta max rel", rel.max(), "exact", int((d["eta"] == r["eta_out"]).sum()), "of", n, flush=True)
    assert rel.max() <= 2e-7
    for k in ("x", "y", "z"):
        ulp = np.abs(d[k].view(np.int32).astype(np.int64) - r[k + "_out"].view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, (k, ulp.max())
    assert np.array_equal(d["tk"] == 0, r["tk_out"] == 0)
    M.close()


def test_movemark_depth_rules_match_reference(gpu):
    """The device movemark step with the depth rules active (g11): types,
    temperatures and positions as the reference's movemark."""
    g = golden("g5_markers.npz")
    r10 = golden("g10_eps_movemark.npz")
    r = golden("g11_movemark_depth.npz")
    print("case: movemark with depth rules", len(g["x"]), flush=True)
    M = _markers(g)
    M.upload(g["x"], g["y"], g["z"], g["tk"], r10["markv"], g["mtype"])
    M.set_strain(r10["marke"])
    dt = float(r["dt"])
    M.update_props([list(x) for x in g["rocks"]], float(g["nubeg"]), float(g["nuend"]), dt, float(r["healrate"]),
                   [r10[k] for k in _ST12], r10["pr"])
    M.set_depth_rules(float(r["zdeep"]), float(r["vdeep"]), float(r["tdeep"]))
    M.advect_rk4(dt, r["vx"], r["vy"], r["vz"])
    d = M.download()
    assert np.array_equal(d["type"], r["type_out"])
    assert np.array_equal(M.get_strain(), r["e_out"])
    for k in ("x", "y", "z"):
        ulp = np.abs(d[k].view(np.int32).astype(np.int64) - r[k + "_out"].view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, (k, ulp.max())
    ok = d["tk"] == r["tk_out"]
    print("tk exact", int(ok.sum()), "of", len(ok), flush=True)
    assert np.array_equal(d["tk"] == 0, r["tk_out"] == 0)
    assert np.allclose(d["tk"], r["tk_out"], rtol=1e-7, atol=0)
    M.set_depth_rules()
    M.close()
