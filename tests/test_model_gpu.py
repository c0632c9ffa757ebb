"""GPU tests of the resident time-step loop (i3g_model_*, the i3mg.c:80-151
sequence vpiterate, titerate, movemark, ronurecalc) on a 33x33x17 grid with
the subduction layout, 2x2x2 markers per cell and Arrhenius viscosity
(E = 300 kJ/mol, capped 1e18-1e25 Pa s).

The loop keeps every field on the device.  Two steps of it equal, to the
bit, the same two steps composed from the host-boundary entry points
(i3g_set_coefficients, i3g_stokes_solve_resident, i3g_stokes_fields,
i3g_heat_solve, i3g_markers_update_tk / advect_rk4 / sort / to_nodes with
host arrays), each of which is checked against the reference elsewhere; and
every step's solve reaches its tolerance, its time step is the displacement
limit, and the markers stay in the domain.
"""
import numpy as np
import pytest

import i3elvis_amd as I

pytestmark = pytest.mark.gpu

ETA_REF = [0, 0, 1e20, 1e23, 1e23, 1e17]
RHO = [3300, 3300, 3250, 3300, 3350, 3250]


def _setup(nx=33, ny=33, nz=17):
    cfg = I.make_config(2, nx, ny, nz, seed=4)
    x, y, z, tk, mt = I.make_markers(cfg, 2, seed=4)
    rocks = I.arrhenius_rocks(ETA_REF, ro=RHO)
    dx = cfg["gx"][1] - cfg["gx"][0]
    p = I.model_params(rtol=1e-8, maxtmstep=3e12, maxxystep=0.5 * dx, heatdif=0.0, nubeg=cfg["nubeg"],
                       nuend=cfg["nuend"])
    return cfg, (x, y, z, tk, mt), rocks, p


def test_loop_equals_host_composition(gpu):
    cfg, mk, rocks, p = _setup()
    x, y, z, tk, mt = mk
    N = cfg["nx"] * cfg["ny"] * cfg["nz"]
    print(f"case: model loop 33x33x17, {len(x)} markers, 2 steps", flush=True)
    Mo = I.Model(cfg, len(x))
    Mo.set_markers(x, y, z, tk, mt, rocks)
    Mo.init(p)
    stats = [Mo.step(p) for _ in range(2)]
    got = Mo.get()
    Mo.close()
    for k, s in enumerate(stats):
        print(f"  step {k}: vcycles {s.vcycles} rel {s.rel_res:.2e} dt {s.timestep:.3e} substeps {s.heat_substeps}"
              f" {s.seconds:.2f}s", flush=True)
        assert s.rel_res <= 1e-8 and s.timestep > 0
    # the same steps through the host-boundary entry points
    S = I.Stokes(cfg)
    H = I.Heat(cfg["gx"], cfg["gy"], cfg["gz"], cfg["bondm"], cfg["bondv"], cfg["bondn"])
    M = I.Markers(cfg["gx"], cfg["gy"], cfg["gz"], len(x))
    M.upload(x, y, z, tk, np.zeros(len(x), np.float32), mt)
    M.set_tbc(cfg["bondm"], cfg["bondv"], cfg["bondn"])
    mkp = dict(nubeg=cfg["nubeg"], nuend=cfg["nuend"], nukoef=0.0, densimod=1, timestep=0.0)
    pr = np.zeros(N)
    M.sort()
    fld = M.to_nodes(rocks, pr, np.zeros((11, N), np.float32), **mkp)
    timesum = 0.0
    for k in range(2):
        S.set_coefficients(nu=fld[0], nxx=fld[7], nxy=fld[8], nxz=fld[9], nyz=fld[10], ro=fld[1])
        st = S.solve(params=p.stokes)
        vx, vy, vz, pr, vmax = S.fields()
        dt = p.maxtmstep
        for v in vmax:
            if v > 0 and dt > p.maxxystep / v:
                dt = p.maxxystep / v
        assert dt == stats[k].timestep and st.vcycles == stats[k].vcycles
        tk0 = fld[6].copy()
        tkn, hs = H.solve(fld[6], fld[1], fld[2], fld[3], fld[4], timestep=dt, maxtkstep=p.heat.maxtkstep,
                          heatmin=p.heat.heatmin, t0koef=p.heat.t0koef, sweeps=p.heat.sweeps,
                          max_sweeps=p.heat.max_sweeps, timesum=timesum, timebeg=p.heat.timebeg,
                          timeend=p.heat.timeend, tempend=p.heat.tempend)
        fld[6] = tkn
        if hs.substeps:
            M.update_tk(fld[6], tk0, fld[3], fld[1], fld[2], p.heatdif, dt)
        M.advect_rk4(dt, vx, vy, vz)
        M.sort()
        fld = M.to_nodes(rocks, pr, fld, **mkp)
        timesum += dt
    for name, a, b in (("vx", got["vx"], vx), ("vy", got["vy"], vy), ("vz", got["vz"], vz), ("pr", got["pr"], pr),
                       ("tk", got["tk"], fld[6]), ("nu", got["nu"], fld[0])):
        assert np.array_equal(a, b), (name, np.abs(a.astype(np.float64) - b).max())
    d = M.download()
    for ax, g in (("x", cfg["gx"]), ("y", cfg["gy"]), ("z", cfg["gz"])):
        assert np.all(d[ax] >= g[0]) and np.all(d[ax] <= g[-1])
    S.close()
    H.close()
    M.close()


def test_loop_learned_start_never_worse_than_previous_solution(gpu):
    """With a learned start each step, the step's start residual is at most
    the previous solution's (the fallback), and the solves converge."""
    import os
    from i3elvis_amd import warmstart as W
    wpath = os.path.join(os.path.dirname(I.__file__), "data", "warmstart_c1_129.npz")
    if not os.path.isfile(wpath):
        pytest.skip("trained weights missing")
    cfg, mk, rocks, p = _setup()
    ws = W.WarmStart.load(wpath, n=None, shape=(cfg["nz"], cfg["nx"], cfg["ny"]))
    p.learned, p.pscale, p.vscale = 1, ws.pscale, ws.vscale
    Mo = I.Model(cfg, len(mk[0]))
    Mo.set_markers(*mk, rocks)
    Mo.init(p)
    for k in range(2):
        s = Mo.step(p, net=ws.net)
        print(f"  step {k}: learned taken {s.learned_taken} vcycles {s.vcycles} rel {s.rel_res:.2e}", flush=True)
        assert s.rel_res <= 1e-8
    Mo.close()
