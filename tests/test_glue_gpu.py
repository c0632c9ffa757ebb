"""GPU tests of the drop-in glue (i3elvis_amd/glue/i3g_glue.c): the
reference's own global state, set up through the reference harness built
with the glue (oracle/_ref/libi3ref_glue.so), is advanced by
i3g_vpiterate / i3g_titerate / i3g_ronurecalc / i3g_movemark, and the
globals are compared with the reference's outputs (golden fixtures) and the
oracle.  This is the drop-in check: same globals in, same globals out."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import golden
from oracle.refh import GLUE_PATH, Ref, available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not available(GLUE_PATH), reason="glue harness not built (needs the reference sources)")]

_R = None


def ref():
    global _R
    if _R is None:
        _R = Ref(GLUE_PATH)
        _R.lib.refh_set_threads(1)
    return _R


def test_glue_vpiterate_matches_reference_converged(gpu):
    """vpiterate through the glue on the reference's globals, solved to a
    relative residual of 1e-10 (i3g_glue_rtol; the default 1e-8 leaves an
    error near 1e-6 that depends on the smoother's path), against the
    reference-assembled direct solve (g3) within 1e-6 per field, with the
    pressure gauge, the CFL time step and the stress fields filled."""
    import i3elvis_amd as I
    n = 17
    print("case: glue vpiterate", n, flush=True)
    g = golden(f"g3_converged_c0_{n}.npz")
    cfg = I.make_config(0, n, n, n, seed=int(g["seed"]))
    R = ref()
    N = n ** 3
    R.setup_grid(cfg["gx"], cfg["gy"], cfg["gz"], cfg["multinum"])
    R.set_bc(cfg["bondm"], cfg["bondv"], cfg["bondn"])
    R.set_physics(g=cfg["g"], nubeg=cfg["nubeg"], nuend=cfg["nuend"])
    R.set("bondnum", len(cfg["bondn"]) - 1, C.c_long)
    for k in ("nu", "ro", "nxx", "nxy", "nxz", "nyz"):
        R.f32(k)[:N] = cfg[k]
    for k in ("vx", "vy", "vz"):
        R.f32(k)[:N] = 0
    R.f64("pr")[:N] = 0
    for k, v in dict(pxinit=2, pyinit=1, pzinit=2).items():
        R.set(k, v, C.c_long)
    R.set("pinit", 0.0)
    R.set("timestep", 1e30)
    R.set("maxxystep", 500.0)
    rtol = R.get("i3g_glue_rtol")
    R.set("i3g_glue_rtol", 1e-10)
    try:
        R.lib.i3g_vpiterate(C.c_int(0), C.c_int(0))
    finally:
        R.set("i3g_glue_rtol", rtol)
    sol = g["sol"].reshape(N, 4)
    ex = g["exists"].reshape(N, 4)
    got = np.stack([R.f64("pr")[:N], R.f32("vx")[:N], R.f32("vy")[:N], R.f32("vz")[:N]], 1)
    for c in range(4):
        e = ex[:, c]
        rel = np.linalg.norm(got[e, c] - sol[e, c]) / np.linalg.norm(sol[e, c])
        print("component", c, "rel L2", rel, flush=True)
        assert rel < 1e-6, (c, rel)
    vmax = max(np.abs(R.f32(k)[:N]).max() for k in ("vx", "vy", "vz"))
    assert R.get("timestep") <= 500.0 / vmax * (1 + 1e-6)
    # the stress fields were filled by the reference's own routines
    out = (C.c_double * 3)()
    R.lib.refh_stress(0, 5, 6, 7, out)
    assert np.isclose(R.f32("sxx", N)[7 * n * n + 5 * n + 6], out[0], rtol=1e-6)


def test_glue_titerate_matches_reference(gpu):
    g = golden("g6_heat.npz")
    n = int(g["n"])
    N = n ** 3
    print("case: glue titerate", n, flush=True)
    R = ref()
    R.setup_grid(g["gx"], g["gy"], g["gz"], 0)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    R.set("bondnum", len(g["bondn"]) - 1, C.c_long)
    for k, v in dict(HEATMIN=1e-8, t0koef=1.0, heatdif=float(g["heatdif"]), timestep=float(g["dt"]),
                     maxtkstep=float(g["maxtkstep"]), timesum=0.0, timebeg=1e30, timeend=2e30, tempend=0.0,
                     zdeep=float(g["zdeep"]), tdeep=float(g["tdeep"])).items():
        R.set(k, v)
    R.i32("multittt", 100)[0] = 10
    m = len(g["x"])
    R.set("marknum", m, C.c_long)
    R.arr("markx", C.c_float, R.MAXMRK)[:m] = g["x"]
    R.arr("marky", C.c_float, R.MAXMRK)[:m] = g["y"]
    R.arr("markz", C.c_float, R.MAXMRK)[:m] = g["z"]
    R.arr("markt", C.c_byte, R.MAXMRK)[:m] = g["mtype"]
    R.arr("markk", C.c_float, R.MAXMRK)[:m] = g["markk"]
    R.arr("markv", C.c_float, R.MAXMRK)[:m] = 0
    for k in ("tk", "ro", "cp", "kt", "ht"):
        R.f32(k)[:N] = g[k]
    R.lib.i3g_titerate(C.c_int(0))
    tk = R.f32("tk")[:N].copy()
    rel = np.linalg.norm(tk.astype(np.float64) - g["tk_out"]) / np.linalg.norm(g["tk_out"])
    print("tk rel L2", rel, flush=True)
    assert rel < 1e-7
    mk = R.arr("markk", C.c_float, R.MAXMRK)[:m]
    assert np.allclose(mk, g["markk_out"], rtol=1e-6, atol=0)


def test_glue_ronurecalc_matches_reference(gpu):
    g = golden("g5_markers.npz")
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    N = nx * ny * nz
    print("case: glue ronurecalc", N, flush=True)
    R = ref()
    R.setup_grid(g["gx"], g["gy"], g["gz"], 0)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    R.set("bondnum", len(g["bondn"]) - 1, C.c_long)
    m = len(g["x"])
    R.set("marknum", m, C.c_long)
    for name, a in (("markx", g["x"]), ("marky", g["y"]), ("markz", g["z"]), ("markk", g["tk"]), ("markv", g["eta"])):
        R.arr(name, C.c_float, R.MAXMRK)[:m] = a
    R.arr("markt", C.c_byte, R.MAXMRK)[:m] = g["mtype"]
    names = ("marknu", "markdh", "markdv", "markn0", "markn1", "markro", "markbb", "markaa", "markcp", "markkt",
             "markkf", "markkp", "markht")
    for q, name in enumerate(names):
        a = R.arr(name, C.c_double, R.MAXTMR)
        a[:] = 0
        a[:len(g["rocks"])] = g["rocks"][:, q]
    R.set("rocknum", len(g["rocks"]), C.c_int)
    R.set("densimod", 1, C.c_int)
    R.set("viscmod", 0, C.c_int)
    for k, v in dict(vdeep=float(g["vdeep"]), nudeep=float(g["nudeep"]), nukoef=0.0, nubeg=float(g["nubeg"]),
                     nuend=float(g["nuend"]), timestep=float(g["timestep"])).items():
        R.set(k, v)
    R.f64("pr")[:N] = g["pr"]
    gn = ["nu", "ro", "cp", "kt", "ht", "et", "tk", "nxx", "nxy", "nxz", "nyz"]
    for q, k in enumerate(gn):
        R.f32(k)[:N] = g["init"][q]
    R.lib.i3g_ronurecalc()
    for q, k in enumerate(gn):
        a = R.f32(k)[:N].astype(np.float64)
        b = g["out"][q].astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        assert rel.max() <= 2e-7, (k, rel.max())
    assert np.allclose(R.arr("markv", C.c_float, R.MAXMRK)[:m], g["eta_out"], rtol=2e-7, atol=0)


def test_glue_movemark_matches_oracle(gpu):
    g = golden("g5_markers.npz")
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    N = nx * ny * nz
    R = ref()
    R.setup_grid(g["gx"], g["gy"], g["gz"], 0)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    R.set("bondnum", len(g["bondn"]) - 1, C.c_long)
    m = len(g["x"])
    R.set("marknum", m, C.c_long)
    for name, a in (("markx", g["x"]), ("marky", g["y"]), ("markz", g["z"]), ("markk", g["tk"]), ("markv", g["eta"])):
        R.arr(name, C.c_float, R.MAXMRK)[:m] = a
    R.arr("markt", C.c_byte, R.MAXMRK)[:m] = g["mtype"]
    R.f32("vx")[:N] = g["vx"]; R.f32("vy")[:N] = g["vy"]; R.f32("vz")[:N] = g["vz"]
    dt = 2e12
    R.set("timestep", dt)
    R.lib.i3g_movemark()
    ox, oy, oz, _, _ = oracle.rk4(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["vx"], g["vy"], g["vz"],
                                  g["gx"], g["gy"], g["gz"], dt)
    for name, b in (("markx", ox), ("marky", oy), ("markz", oz)):
        a = R.arr(name, C.c_float, R.MAXMRK)[:m]
        ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, (name, ulp.max())


def test_glue_titerate_with_heating_terms_matches_reference(gpu):
    """i3g_titerate on the reference's globals with adiabyn 2 and frictyn 1
    (g8, written from the reference's own titerate): the glue passes
    vx/vy/vz, pr, et and the stress/strain-rate fields to the device heat
    step, which reproduces the reference temperature."""
    g = golden("g8_heat_sources.npz")
    n = int(g["n"])
    N = n ** 3
    print("case: glue titerate with heating terms", n, flush=True)
    R = ref()
    R.setup_grid(g["gx"], g["gy"], g["gz"], 0)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    R.set("bondnum", len(g["bondn"]) - 1, C.c_long)
    for k, v in dict(HEATMIN=1e-8, t0koef=1.0, heatdif=0.0, timestep=float(g["dt"]), maxtkstep=float(g["maxtkstep"]),
                     timesum=0.0, timebeg=1e30, timeend=2e30, tempend=0.0, zdeep=1e30, tdeep=0.0,
                     ysize=float(g["ysize"]), stp100=float(g["stp100"]), GXKOEF=float(g["gk"][0]),
                     GYKOEF=float(g["gk"][1]), GZKOEF=float(g["gk"][2])).items():
        R.set(k, v)
    R.set("adiabyn", 2, C.c_int)
    R.set("frictyn", 1, C.c_int)
    R.i32("multittt", 100)[0] = 10
    # a few markers inside the grid (their update follows the solve)
    rng = np.random.default_rng(3)
    m = 64
    R.set("marknum", m, C.c_long)
    for k, gl in (("markx", g["gx"]), ("marky", g["gy"]), ("markz", g["gz"])):
        R.arr(k, C.c_float, R.MAXMRK)[:m] = rng.uniform(gl[1], gl[-2], m)
    R.arr("markt", C.c_byte, R.MAXMRK)[:m] = 2
    R.arr("markk", C.c_float, R.MAXMRK)[:m] = 1000.0
    R.arr("markv", C.c_float, R.MAXMRK)[:m] = 0
    for k in ("tk", "ro", "cp", "kt", "ht", "et", "vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz",
              "exx", "eyy", "ezz", "exy", "exz", "eyz"):
        R.f32(k)[:N] = g[k]
    R.f64("pr")[:N] = g["pr"]
    R.lib.i3g_titerate(C.c_int(0))
    R.set("adiabyn", 0, C.c_int)
    R.set("frictyn", 0, C.c_int)
    tk = R.f32("tk")[:N].copy()
    rel = np.linalg.norm(tk.astype(np.float64) - g["tk_out"]) / np.linalg.norm(g["tk_out"])
    print("tk rel L2", rel, flush=True)
    assert rel < 1e-7


@pytest.mark.parametrize("case", ["g10", "g11"])
def test_glue_movemark_matches_reference(gpu, case):
    """i3g_movemark on the reference's globals (g10 inputs: the g5 markers
    with random accumulated strain, stress/strain-rate, pressure and
    velocity fields; g11: faster flow with the zdeep/vdeep depth rules
    active): viscosity and strain from epscalc1 + viscalc at the old
    positions, then RK4 with the depth rules, against the reference's own
    movemark."""
    g = golden("g5_markers.npz")
    r10 = golden("g10_eps_movemark.npz")
    r = dict(r10)
    depth = dict(zdeep=1e30, vdeep=1e30, tdeep=0.0)
    if case == "g11":
        r11 = golden("g11_movemark_depth.npz")
        r.update(dict(r11))
        depth = dict(zdeep=float(r11["zdeep"]), vdeep=float(r11["vdeep"]), tdeep=float(r11["tdeep"]))
    nx, ny, nz = len(g["gx"]), len(g["gy"]), len(g["gz"])
    N = nx * ny * nz
    print("case: glue movemark vs reference", len(g["x"]), flush=True)
    R = ref()
    R.setup_grid(g["gx"], g["gy"], g["gz"], 0)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    R.set("bondnum", len(g["bondn"]) - 1, C.c_long)
    m = len(g["x"])
    R.set("marknum", m, C.c_long)
    for name, a in (("markx", g["x"]), ("marky", g["y"]), ("markz", g["z"]), ("markk", g["tk"]), ("markv", r["markv"]),
                    ("marke", r["marke"])):
        R.arr(name, C.c_float, R.MAXMRK)[:m] = a
    R.arr("markt", C.c_byte, R.MAXMRK)[:m] = g["mtype"]
    names = dict(nu0="marknu", dh="markdh", dv="markdv", n0="markn0", n1="markn1", ro="markro", bro="markbb",
                 aro="markaa", cp="markcp", kt="markkt", kf="markkf", kp="markkp", ht="markht")
    fields = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")
    for q, f in enumerate(fields):
        a = R.arr(names[f], C.c_double, R.MAXTMR)
        a[:] = 0
        a[:len(g["rocks"])] = g["rocks"][:, q]
    R.set("rocknum", len(g["rocks"]), C.c_int)
    for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz", "vx", "vy", "vz"):
        R.f32(k)[:N] = r[k]
    R.f64("pr")[:N] = r["pr"]
    for k, v in dict(timestep=float(r["dt"]), healrate=float(r["healrate"]), nubeg=float(g["nubeg"]),
                     nuend=float(g["nuend"]), nudeep=1e30, nukoef=0.0, **depth).items():
        R.set(k, v)
    R.set("viscmod", 0, C.c_int)
    R.set("densimod", 1, C.c_int)
    R.lib.i3g_movemark()
    eta = R.arr("markv", C.c_float, R.MAXMRK)[:m]
    rel = np.abs(eta.astype(np.float64) - r["eta_out"]) / np.maximum(np.abs(r["eta_out"]), 1e-30)
    print("eta max rel", rel.max(), flush=True)
    assert rel.max() <= 2e-7
    assert np.array_equal(R.arr("marke", C.c_float, R.MAXMRK)[:m], r["e_out"])
    for name, k in (("markx", "x_out"), ("marky", "y_out"), ("markz", "z_out")):
        a = R.arr(name, C.c_float, R.MAXMRK)[:m]
        ulp = np.abs(a.view(np.int32).astype(np.int64) - r[k].view(np.int32).astype(np.int64))
        assert ulp.max() <= 1, (name, ulp.max())
    assert np.array_equal(R.arr("markk", C.c_float, R.MAXMRK)[:m] == 0, r["tk_out"] == 0)
    assert np.array_equal(R.arr("markt", C.c_byte, R.MAXMRK)[:m], r["type_out"])
