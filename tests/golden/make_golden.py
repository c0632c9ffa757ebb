"""Writes the golden fixtures under tests/golden/ from the reference itself.

Runs only where the reference checkout exists: it loads the reference
harness oracle/_ref/libi3ref.so (the reference's solver sources compiled by
oracle/Makefile), sets the reference's globals and calls its routines with
one OpenMP thread (its scatter loops race otherwise).  The fixtures are
inputs and outputs only.

  g1_operator.npz   level-0/1 residuals b - A x of the reference's assembled
                    rows on a random non-uniform grid with random
                    log-uniform viscosities; ronurestrict, vprestrict and
                    vpprolong outputs on random vectors (move3mg.c:718-1579,
                    2642-3520)
  g3_converged_*.npz  converged FP64 solution of the reference's assembled
                    level-0 system (configuration 0 geometry; c2: the
                    subduction layout) by an equilibrated sparse direct
                    solve, pressure gauge P(2,1,2) = 0 as vpiterate
                    (move3mg.c:2248-2269)
  g13_*             the reference's run files: g13_mode.t3c / g13_file.t3c (a
                    deck of this build's own values in the reference's
                    format) and what the reference's loadconf reads from it;
                    g13_state.prn written by the reference's saver from the
                    state in g13_io.npz (load3mg.c:2-180, 704-1096)
  g14_movemark_rheology.npz  movemark with the deck's rock table (init.t3c):
                    full viscalc (brittle, power law, Karato & Wu, grain
                    size), outgrid 1/2, markmod 0/1, fluid and immobile
                    markers (mark3mg.c:2-397, 1114-1681)
  g15_ronurecalc_rheology.npz  ronurecalc with the deck's rocks: retyping,
                    water/air mixing, viscalc from strain rates, phase
                    densities (mark3mg.c:404-1015)
  g16_*             the reference's generator in3mg run on g16_init.t3c (a
                    deck of this build's own values) with g16_mode.t3c: the
                    state it saves (g16_generator.npz) and its ParaView
                    nodal and composition files (in3mg.c:548-1049,
                    i3vis_output_paraview.c)
  g12_smoother.npz  on the g1 inputs: the reference's own coefficient of each
                    assembled row's unknown (the first row entry, gausmat4's
                    divisor ival1, move3mg.c:3096-3124) and the state after
                    one gausmat4 sweep on level 0 from g1's x0 (one thread,
                    lexicographic, move3mg.c:3087-3131)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import i3elvis_amd as I  # noqa: E402  (host-side generators only)
from oracle.refh import Ref  # noqa: E402


def reference_system(R, l):
    import scipy.sparse as sp
    n4 = int(R.mgn[l]) * 4
    rowptr, cols, vals, rhs = R.assemble_level(l)
    A = sp.csr_matrix((vals, cols - 4 * int(R.mgp[l]), rowptr), shape=(n4, n4))
    return A, rhs, np.diff(rowptr) > 0


def g1(R):
    rng = np.random.default_rng(20240601)
    nx, ny, nz, mn = 13, 11, 12, 1
    gx = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, nx - 1)]) * 2e3
    gy = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, ny - 1)]) * 2e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, nz - 1)]) * 2e3
    N0 = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, mn)
    rules, nr = I.freeslip_rules()
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    f = {k: (10 ** rng.uniform(19, 23, N0)).astype(np.float32) for k in ("nu", "nxx", "nxy", "nxz", "nyz")}
    f["ro"] = rng.uniform(3000, 3400, N0).astype(np.float32)
    R.setup_grid(gx, gy, gz, mn)
    R.set_bc(bondm, bondv, bondn)
    R.set_physics(g=(0.3, 9.81, -0.2), nubeg=1e19, nuend=1e23)
    for k, v in f.items():
        R.f32(k)[:N0] = v
    A0, rhs0, ex0 = reference_system(R, 0)
    x0 = rng.standard_normal(4 * N0) * np.tile([1e7, 1e-9, 1e-9, 1e-9], N0)
    x0[~ex0] = 0
    r0 = rhs0 - A0 @ x0
    R.lib.refh_ronurestrict(1)
    p1, N1 = int(R.mgp[1]), int(R.mgn[1])
    nu1 = R.f32("nu")[p1:p1 + N1].copy()
    A1, rhs1, ex1 = reference_system(R, 1)
    x1 = rng.standard_normal(4 * N1) * np.tile([1e7, 1e-9, 1e-9, 1e-9], N1)
    x1[~ex1] = 0
    b1 = rng.standard_normal(4 * N1)
    b1[~ex1] = 0
    r1 = b1 - A1 @ x1
    nd = R.f32("nd")
    nd0, nd1 = nd[:N0].copy(), nd[p1:p1 + N1].copy()
    rr = rng.standard_normal(4 * N0)
    R.f64("fre0", R.MAXPPP)[:4 * N0] = rr
    R.lib.refh_vprestrict(1)
    restr = R.f64("fre1", R.MAXPPP)[4 * p1:4 * (p1 + N1)].copy()
    sol0 = R.f64("sol0", R.MAXPPP)
    xc = rng.standard_normal(4 * N1)
    xf = rng.standard_normal(4 * N0)
    sol0[:4 * N0] = xf
    sol0[4 * p1:4 * (p1 + N1)] = xc
    R.lib.refh_vpprolong(1)
    prol = sol0[:4 * N0].copy()
    lines1 = R.level_lines(1)
    np.savez_compressed(
        os.path.join(HERE, "g1_operator.npz"), nx=nx, ny=ny, nz=nz, multinum=mn, gx=gx, gy=gy, gz=gz,
        g=np.array([0.3, 9.81, -0.2]), nubeg=1e19, nuend=1e23, bondm=bondm, bondv=bondv, bondn=bondn,
        **f, x0=x0, rhs0=rhs0, r0=r0, exists0=ex0, nu1=nu1, x1=x1, b1=b1, r1=r1, exists1=ex1,
        lx1=lines1[0], ly1=lines1[1], lz1=lines1[2], mgx=R.mgx, mgy=R.mgy, mgz=R.mgz, mgp=R.mgp,
        nd0=nd0, nd1=nd1, restr_in=rr, restr_out=restr, prol_xc=xc, prol_xf=xf, prol_out=prol,
        v1koef=0.7, p1koef=0.1)
    print("g1_operator.npz written")


def g3(R, n, which=0, seed=1):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    cfg = I.make_config(which, n, n, n, seed=seed)
    N0 = n ** 3
    R.setup_grid(cfg["gx"], cfg["gy"], cfg["gz"], cfg["multinum"])
    R.set_bc(cfg["bondm"], cfg["bondv"], cfg["bondn"])
    R.set_physics(g=cfg["g"], nubeg=cfg["nubeg"], nuend=cfg["nuend"])
    for k in ("nu", "ro", "nxx", "nxy", "nxz", "nyz"):
        R.f32(k)[:N0] = cfg[k]
    A, rhs, ex = reference_system(R, 0)
    A = A.tolil()
    for u in np.where(~ex)[0]:
        A[u, u] = 1.0
    ref = (2 * n * n + 2 * n + 1) * 4  # P at (m1,m2,m3) = (2,1,2)
    A[ref, :] = 0
    A[ref, ref] = 1.0
    b = rhs.copy()
    b[ref] = 0
    A = A.tocsr()
    comp = np.arange(4 * N0) % 4
    D = np.where(comp == 0, 1e7, 1e-9)
    As = A @ sp.diags(D)
    Rs = 1.0 / abs(As).max(axis=1).toarray().ravel()
    As = sp.diags(Rs) @ As
    lu = spl.splu(As.tocsc())
    y = lu.solve(Rs * b)
    # iterative refinement on the equilibrated system: the 10^6-contrast
    # subduction system leaves the first solve's pressure at ~1e-6 relative
    for _ in range(4):
        y = y + lu.solve(Rs * b - As @ y)
    x = D * y
    r = rhs - (A @ x)
    r[ref] = 0
    np.savez_compressed(os.path.join(HERE, f"g3_converged_c{which}_{n}.npz"), n=n, which=which, seed=seed,
                        sol=x, rhs=rhs, exists=ex, resid_max=np.abs(r).max())
    print(f"g3_converged_c{which}_{n}.npz written; max |b - A x| = {np.abs(r).max():.3e}")


def g12(R):
    g = dict(np.load(os.path.join(HERE, "g1_operator.npz")))
    nx, ny, nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    N0 = nx * ny * nz
    R.setup_grid(g["gx"], g["gy"], g["gz"], int(g["multinum"]))
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    v0koef, p0koef = 0.7, 0.3
    R.set_physics(g=tuple(g["g"]), nubeg=float(g["nubeg"]), nuend=float(g["nuend"]), v0koef=v0koef, p0koef=p0koef)
    for k in ("nu", "nxx", "nxy", "nxz", "nyz", "ro"):
        R.f32(k)[:N0] = g[k]
    rowptr, cols, vals, rhs = R.assemble_level(0)
    assert np.array_equal(rhs, g["rhs0"])
    has = np.diff(rowptr) > 0
    # velocity and BC rows start with their own unknown; continuity rows
    # have none (no P diagonal) and use nd * p0koef instead
    vel = has & ((np.arange(4 * N0) % 4 != 0) | (g["bondm"][:4 * N0] != 0))
    diag = np.zeros(4 * N0)
    diag[vel] = vals[rowptr[:-1][vel]]
    assert np.array_equal(cols[rowptr[:-1][vel]], np.arange(4 * N0)[vel]), "first row entry is not the row's unknown"
    R.lib.refh_ronurestrict(1)
    R.lib.refh_store_all_levels()
    sol0 = R.f64("sol0", R.MAXPPP)
    sol0[:] = 0
    sol0[:4 * N0] = g["x0"]
    R.lib.refh_gausmat4(1, 0)
    gs_out = sol0[:4 * N0].copy()
    nd0 = R.f32("nd")[:N0].copy()
    np.savez_compressed(os.path.join(HERE, "g12_smoother.npz"), diag0=diag, gs_out=gs_out, nd0=nd0,
                        v0koef=v0koef, p0koef=p0koef)
    print("g12_smoother.npz written")


DECK = """/i3g_test_deck_in_the_reference_layout_(own_values)
run/start.prn b

/SAVEFILE_TYPE_cyc0max_maxxystep_maxtkstep_maxtmstep(yr)_seven_more_steps_three_cycle_counts
out_a.prn b 4 2000.0 50.0 1e4  0.5 1.5 2.5 3.5 4.5 5.5 6.5 7.5 10 20 30
out_b.prn t 2 1000.0 25.0 2e4  0.1 0.2 0.3 0.4 0.5 0.6 0.7 0.8 11 21 31
/a_comment_between_entries
out_c.prn b 1 500.0 10.0 3.5e3 9 8 7 6 5 4 3 2 12 22 32
~
/GENERAL
0-printmod
1-movemod
1-tempmod
1-markmod
1-gridmod
2-outgrid
1-densimod
1e3-stp100
4-filestop
/EROSION
1.1e4-eroslev 1.2e4-sedilev 1.3e4-waterlev
/V_P
100-cyc1max
1e-11-DIVVMIN
2e-10-STOKSMIN
3e-9-DIVVMAX
4e-8-STOKSMAX
3-multinum
4-multicyc
8 16 16 32
8 16 16 64
0.31-p0koef 0.12 0.0 0.71 0.69 1e18 1e25 0.001 1-viscmod 0-spheryn
/T
7-cyc2max 1e-5-HEATMIN 0-multinumt 3-multicyct
30
31
0.9 0.8 1e-12-heatdif 1-frictyn 2.7-adiabyn
/HYDRATION_AND_THE_REST
373 0.1 -1e-9-vyfluid -2e-9 0.05 1573 0.5 0.2 6.6e5 6.0e5 1e20 1500 1600 1700 0.1 2e4 873 0.9 1.2e4 3e3
2 1e4 8e2 3e-14 0.01 0.02 0.03 1073 1273 0.4 0.6
/PLUME
1e6-timebeg 2e6-timeend 1800-tempbeg 5e6-convend
"""


def g13(R):
    import shutil
    import tempfile
    import ctypes as C
    with open(os.path.join(HERE, "g13_mode.t3c"), "w") as f:
        f.write(DECK)
    with open(os.path.join(HERE, "g13_file.t3c"), "w") as f:
        f.write("2 \n")
    tmp = tempfile.mkdtemp()
    cwd = os.getcwd()
    try:
        shutil.copy(os.path.join(HERE, "g13_mode.t3c"), os.path.join(tmp, "mode.t3c"))
        shutil.copy(os.path.join(HERE, "g13_file.t3c"), os.path.join(tmp, "file.t3c"))
        os.chdir(tmp)
        fln3 = R.lib.refh_loadconf()
        gi = lambda n: C.c_int.in_dll(R.lib, n).value
        gd = lambda n: C.c_double.in_dll(R.lib, n).value
        fl0num = gi("fl0num")
        chars = lambda n, k: bytes((C.c_char * 60).in_dll(R.lib, n)).split(b"\0", 1)[0].decode()
        fl0out = [bytes(np.ctypeslib.as_array((C.c_char * (60 * 5000)).in_dll(R.lib, "fl0out"))[60 * k:60 * k + 60])
                  .split(b"\0", 1)[0].decode() for k in range(fl0num)]
        fl0otp = np.ctypeslib.as_array((C.c_int * 5000).in_dll(R.lib, "fl0otp"))[:fl0num].copy()
        fl0stp = np.ctypeslib.as_array((C.c_double * (5000 * 20)).in_dll(R.lib, "fl0stp")).reshape(5000, 20)[:fl0num, :11].copy()
        fl0cyc = np.ctypeslib.as_array((C.c_int * (5000 * 5)).in_dll(R.lib, "fl0cyc")).reshape(5000, 5)[:fl0num, :4].copy()
        mn, mnt = gi("multinum"), gi("multinumt")
        mnnn = np.ctypeslib.as_array((C.c_int * (5 * 50)).in_dll(R.lib, "multinnn")).reshape(5, 50)[:2, :mn + 1].copy()
        mttt = np.ctypeslib.as_array((C.c_int * (2 * 50)).in_dll(R.lib, "multittt")).reshape(2, 50)[:2, :mnt + 1].copy()
        ints = {n: gi(n) for n in ("fl1itp", "printmod", "movemod", "tempmod", "markmod", "gridmod", "outgrid",
                                    "densimod", "filesjob", "cyc1max", "multinum", "multicyc", "viscmod", "spheryn",
                                    "cyc2max", "multinumt", "multicyct", "frictyn", "adiabyn")}
        dbls = {n: gd(n) for n in ("stp100", "eroslev", "sedilev", "waterlev", "DIVVMIN", "STOKSMIN", "DIVVMAX",
                                    "STOKSMAX", "p0koef", "p1koef", "p2koef", "v0koef", "v1koef", "nubeg", "nuend",
                                    "nukoef", "HEATMIN", "t0koef", "t1koef", "heatdif", "tkpor", "zmpor", "vyfluid",
                                    "vymelt", "dmwamin", "tdeep", "dtdeep", "drdeep", "zdeep", "vdeep", "nudeep",
                                    "dxwater", "dywater", "dzwater", "dserp", "yserp", "tserp", "lambmlt", "ydike",
                                    "dydike", "nusseltno", "circdepth", "circtemp", "healrate", "meltmin",
                                    "maxwater", "maxpmelt", "teclmin", "teclmax", "xvolc", "xtkplut", "timebeg",
                                    "timeend", "tempbeg", "convend")}
        fl1in = chars("fl1in", 0)
        # the .prn state: a small grid, rocks, BC records and markers
        rng = np.random.default_rng(1313)
        nx, ny, nz, nr, nm = 5, 4, 3, 3, 7
        N = nx * ny * nz
        gxv = np.cumsum(rng.uniform(1, 2, nx)) * 1e4
        gyv = np.cumsum(rng.uniform(1, 2, ny)) * 1e4
        gzv = np.cumsum(rng.uniform(1, 2, nz)) * 1e4
        R.setup_grid(gxv, gyv, gzv, 0)  # gridcheck: nodenum and the level tables
        st = {}
        for nme, ct, v in (("xnumx", C.c_long, nx), ("ynumy", C.c_long, ny), ("znumz", C.c_long, nz),
                           ("mnumx", C.c_long, 2), ("mnumy", C.c_long, 3), ("mnumz", C.c_long, 4),
                           ("pxinit", C.c_long, 2), ("pyinit", C.c_long, 1), ("pzinit", C.c_long, 2),
                           ("rocknum", C.c_int, nr), ("marknum", C.c_long, nm)):
            ct.in_dll(R.lib, nme).value = v
            st[nme] = v
        for nme, v in (("xsize", 4e5), ("ysize", 3e5), ("zsize", 2e5), ("pinit", 1e5), ("GXKOEF", 0.1),
                       ("GYKOEF", 9.81), ("GZKOEF", -0.2), ("timesum", 3.15576e7 * 1234.5), ("gridcur", 0.25),
                       ("gridtot", 0.75)):
            C.c_double.in_dll(R.lib, nme).value = v
            st[nme] = v
        rock_names = ("markn0", "markn1", "marks0", "marks1", "marknu", "markdh", "markdv", "markss", "markmm",
                      "markll", "marka0", "marka1", "markb0", "markb1", "marke0", "marke1", "markro", "markbb",
                      "markaa", "markcp", "markkt", "markkf", "markkp", "markht")
        rock = rng.uniform(-1, 1, (24, nr)) * 10.0 ** rng.integers(-8, 20, (24, 1))
        for c, nme in enumerate(rock_names):
            R.arr(nme, C.c_double, nr)[:] = rock[c]
        node_names = ("pr", "vx", "vy", "vz", "ro", "nu", "nxx", "nxy", "nxz", "nyz", "tk", "cp", "et", "kt", "ht",
                      "rr")
        node = (rng.standard_normal((16, N)) * 10.0 ** rng.integers(-9, 9, (16, 1))).astype(np.float32)
        node[0] = rng.standard_normal(N) * 1e8  # pr: double in the program, float in the file
        for c, nme in enumerate(node_names):
            if nme == "pr":
                R.f64("pr")[:N] = node[0].astype(np.float64) + 0.123  # rounded to float by saver
            else:
                R.f32(nme)[:N] = node[c]
        node[0] = (node[0].astype(np.float64) + 0.123).astype(np.float32)
        bm = R.i32("bondm", 10 * N)
        bm[:] = 0
        at = np.sort(rng.choice(10 * N, 9, replace=False))
        bm[at] = np.arange(1, 10)
        bv = R.arr("bondv1", C.c_float, 2 * 20)
        bv[:] = rng.standard_normal(40).astype(np.float32)
        bn = R.i32("bondn1", 20)
        bn[:] = rng.integers(0, 4 * N, 20)
        C.c_long.in_dll(R.lib, "bondnum").value = 9
        mark_names = ("markx", "marky", "markz", "markk", "markw", "markd", "markv", "marke", "markex", "markft",
                      "markrr", "markrn")
        mark = rng.standard_normal((12, nm)).astype(np.float32) * 1e3
        for c, nme in enumerate(mark_names):
            R.f32(nme, R.MAXMRK)[:nm] = mark[c]
        mt = rng.integers(0, 30, nm).astype(np.int8)
        np.ctypeslib.as_array((C.c_char * R.MAXMRK).in_dll(R.lib, "markt")).view(np.int8)[:nm] = mt
        with open("stop.yn", "w") as f:
            f.write("n \n")
        R.lib.refh_saver(b"state.prn", 0, 6)
        shutil.copy("state.prn", os.path.join(HERE, "g13_state.prn"))
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp)
    np.savez_compressed(os.path.join(HERE, "g13_io.npz"), fln3=fln3, fl1in=np.array(fl1in), fl0num=fl0num,
                        fl0out=np.array(fl0out), fl0otp=fl0otp, fl0stp=fl0stp, fl0cyc=fl0cyc, multinnn=mnnn,
                        multittt=mttt, int_names=np.array(list(ints)), int_vals=np.array(list(ints.values())),
                        dbl_names=np.array(list(dbls)), dbl_vals=np.array(list(dbls.values())),
                        prn_rock=rock, prn_node=node, prn_gx=gxv.astype(np.float32), prn_gy=gyv.astype(np.float32),
                        prn_gz=gzv.astype(np.float32), prn_bond_at=at, prn_bond_rec=np.arange(1, 10),
                        prn_bond_v=np.array([[bv[2 * r], bv[2 * r + 1]] for r in range(1, 10)], np.float32),
                        prn_bond_partner=np.array([bn[r] for r in range(1, 10)]), prn_mark=mark, prn_markt=mt,
                        prn_n0=6, **{"prn_" + k: v for k, v in st.items()})
    print("g13 written: fl0num", fl0num, "fln3", fln3, "fl1in", fl1in)


ROCKS = [  # types 0,1 unused (air/water branches), 2..6 used
    dict(nu0=1e21, n0=1e18, n1=1e25, ro=3000, cp=1000, kt=2.0),
    dict(nu0=1e21, n0=1e18, n1=1e25, ro=1000, cp=1000, kt=2.0),
    dict(nu0=1e21, dh=0, dv=0, n0=1e18, n1=1e25, ro=3300, bro=3e-5, aro=1e-3, cp=1000, kt=3.0, kf=0, kp=0, ht=1e-9),
    dict(nu0=1e23, dh=0, dv=0, n0=1e18, n1=1e25, ro=3200, bro=2e-5, aro=1e-3, cp=1100, kt=0.64, kf=807.0, kp=4e-6, ht=2e-6),
    dict(nu0=2e16, dh=3e5, dv=4.0, n0=1e19, n1=1e24, ro=3350, bro=3e-5, aro=1e-3, cp=1000, kt=0.73, kf=1293.0, kp=4e-6, ht=2e-8),
    dict(nu0=1e10, dh=3e5, dv=0, n0=1e18, n1=1e25, ro=3250, bro=3e-5, aro=1e-3, cp=1000, kt=1.0, kf=500.0, kp=0, ht=0),
    dict(nu0=0, dh=0, dv=0, n0=1e20, n1=5e22, ro=2900, bro=3e-5, aro=1e-3, cp=1200, kt=2.5, kf=0, kp=0, ht=1e-7),
]


def g5(R):
    """Marker -> node (ronurecalc), cell search and vxyzcalc1 through the
    reference, on a small non-uniform grid."""
    import ctypes as C
    rng = np.random.default_rng(77)
    nx, ny, nz = 11, 9, 10
    gx = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nx - 1)]) * 5e3
    gy = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, ny - 1)]) * 5e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nz - 1)]) * 5e3
    N = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, 0)
    rules, nr = I.freeslip_rules(273.0, 1600.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(bondm, bondv, bondn)
    # markers: 2x2x2 per cell with jitter, generation order = cell order,
    # plus a few outside the domain and on cell lines
    xs, ys, zs = [], [], []
    for k in range(nz - 1):
        for i in range(nx - 1):
            for j in range(ny - 1):
                for a in range(2):
                    for b in range(2):
                        for c in range(2):
                            xs.append(gx[i] + (b + 0.5 + rng.uniform(-0.45, 0.45)) * (gx[i + 1] - gx[i]) / 2)
                            ys.append(gy[j] + (c + 0.5 + rng.uniform(-0.45, 0.45)) * (gy[j + 1] - gy[j]) / 2)
                            zs.append(gz[k] + (a + 0.5 + rng.uniform(-0.45, 0.45)) * (gz[k + 1] - gz[k]) / 2)
    extra = 64
    xs += list(rng.uniform(-2e3, gx[-1] + 2e3, extra))
    ys += list(rng.uniform(-2e3, gy[-1] + 2e3, extra))
    zs += list(rng.uniform(-2e3, gz[-1] + 2e3, extra))
    x = np.array(xs, np.float32); y = np.array(ys, np.float32); z = np.array(zs, np.float32)
    x[:5] = np.float32(gx[3]); y[5:10] = np.float32(gy[4])  # exactly on grid lines
    n = len(x)
    mtype = rng.integers(2, 7, n).astype(np.int8)
    tk = rng.uniform(600, 1700, n).astype(np.float32)
    tk[::97] = 0  # excluded markers (markk <= 0)
    eta = np.where(rng.uniform(size=n) < 0.5, 0.0, 10 ** rng.uniform(19, 23, n)).astype(np.float32)
    pr = rng.uniform(1e7, 1e9, N)
    # reference globals
    R.set("marknum", n, C.c_long)
    R.arr("markx", C.c_float, R.MAXMRK)[:n] = x
    R.arr("marky", C.c_float, R.MAXMRK)[:n] = y
    R.arr("markz", C.c_float, R.MAXMRK)[:n] = z
    R.arr("markk", C.c_float, R.MAXMRK)[:n] = tk
    R.arr("markv", C.c_float, R.MAXMRK)[:n] = eta
    R.arr("markt", C.c_byte, R.MAXMRK)[:n] = mtype
    R.arr("markrr", C.c_float, R.MAXMRK)[:n] = 0
    R.arr("marke", C.c_float, R.MAXMRK)[:n] = 0
    names = dict(nu0="marknu", dh="markdh", dv="markdv", n0="markn0", n1="markn1", ro="markro", bro="markbb",
                 aro="markaa", cp="markcp", kt="markkt", kf="markkf", kp="markkp", ht="markht")
    for f, gname in names.items():
        a = R.arr(gname, C.c_double, R.MAXTMR)
        a[:] = 0
        for t, r in enumerate(ROCKS):
            a[t] = r.get(f, 0.0)
    for gname in ("markss", "markll", "marka0", "marka1", "markb0", "markb1", "marke0", "marke1"):
        R.arr(gname, C.c_double, R.MAXTMR)[:] = 0
    R.arr("markmm", C.c_double, R.MAXTMR)[:] = 1.0
    R.set("densimod", 1, C.c_int)
    for k, v in dict(eroslev=-1e30, sedilev=1e30, waterlev=1e30, vdeep=gy[-1] * 0.8, nudeep=1e21, nukoef=0.0,
                     nubeg=1e18, nuend=1e25, timestep=1e10).items():
        R.set(k, v)
    R.set("viscmod", 0, C.c_int)
    R.f64("pr")[:N] = pr
    init = np.zeros((11, N), np.float32)
    init[0] = 5e20; init[1] = 3111; init[2] = 999; init[3] = 2.2; init[4] = 1e-9; init[5] = 1e-5
    init[6] = 1000; init[7:] = 7e20
    gnames = ["nu", "ro", "cp", "kt", "ht", "et", "tk", "nxx", "nxy", "nxz", "nyz"]
    for q, gname in enumerate(gnames):
        R.f32(gname)[:N] = init[q]
    R.lib.refh_ronurecalc()
    out = np.stack([R.f32(gname)[:N].copy() for gname in gnames])
    eta_out = R.arr("markv", C.c_float, R.MAXMRK)[:n].copy()
    # cell search at random points (clamped as prcalc1 does)
    pts = np.stack([rng.uniform(-3e3, gx[-1] + 3e3, 400), rng.uniform(-3e3, gy[-1] + 3e3, 400),
                    rng.uniform(-3e3, gz[-1] + 3e3, 400)], 1).astype(np.float32)
    pts[:20, 0] = gx[rng.integers(0, nx, 20)]
    cells = np.array([[R.lib.refh_m1serch(min(max(float(p[0]), 0.0), gx[-1])),
                       R.lib.refh_m2serch(min(max(float(p[1]), 0.0), gy[-1])),
                       R.lib.refh_m3serch(min(max(float(p[2]), 0.0), gz[-1]))] for p in pts], np.int32)
    # staggered velocity interpolation
    vxf = rng.standard_normal(N).astype(np.float32) * 1e-9
    vyf = rng.standard_normal(N).astype(np.float32) * 1e-9
    vzf = rng.standard_normal(N).astype(np.float32) * 1e-9
    R.f32("vx")[:N] = vxf; R.f32("vy")[:N] = vyf; R.f32("vz")[:N] = vzf
    vout = np.zeros((len(pts), 3))
    buf = (C.c_double * 3)()
    for q, p in enumerate(pts):
        R.lib.refh_vxyzcalc1(C.c_double(float(p[0])), C.c_double(float(p[1])), C.c_double(float(p[2])), buf)
        vout[q] = list(buf)
    np.savez_compressed(os.path.join(HERE, "g5_markers.npz"), gx=gx, gy=gy, gz=gz, bondm=bondm, bondv=bondv,
                        bondn=bondn, x=x, y=y, z=z, mtype=mtype, tk=tk, eta=eta, pr=pr, init=init, out=out,
                        eta_out=eta_out, pts=pts, cells=cells, vx=vxf, vy=vyf, vz=vzf, vout=vout,
                        rocks=np.array([[r.get(f, 0.0) for f in ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro",
                                                                    "cp", "kt", "kf", "kp", "ht")] for r in ROCKS]),
                        nubeg=1e18, nuend=1e25, vdeep=gy[-1] * 0.8, nudeep=1e21, timestep=1e10)
    print("g5_markers.npz written")


def g6(R, n=17):
    """Implicit heat step with sub-steps (titerate, heat3mg.c:287-588) and
    the marker temperature update with subgrid diffusion (heat3mg.c:592-782)
    through the reference, configuration-0 geometry."""
    import ctypes as C
    import oracle
    rng = np.random.default_rng(606)
    cfg = I.make_config(0, n, n, n, seed=1)
    gx, gy, gz = cfg["gx"], cfg["gy"], cfg["gz"]
    N = n ** 3
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(cfg["bondm"], cfg["bondv"], cfg["bondn"])
    ro = cfg["ro"].astype(np.float32)
    cp = rng.uniform(900, 1300, N).astype(np.float32)
    kt = rng.uniform(1.5, 4.5, N).astype(np.float32)
    ht = rng.uniform(0, 2e-6, N).astype(np.float32)
    tk = (cfg["tk"] + rng.uniform(-50, 50, N)).astype(np.float32)
    dt = 1e13
    H = oracle.heat_sys(gx, gy, gz, ro, cp, kt, ht, tk, dt, cfg["bondm"], cfg["bondv"], cfg["bondn"])
    maxtkstep = oracle.heat_maxrate(H) * dt / 2.5  # three sub-steps
    # markers: 2x2x2 per cell with jitter, a few outside the grid or of type >= 50
    cells = np.stack(np.meshgrid(np.arange(n - 1), np.arange(n - 1), np.arange(n - 1), indexing="ij"), -1).reshape(-1, 3)
    c = np.repeat(cells, 8, 0)
    def jit(g, k):
        return g[k] + rng.uniform(0.02, 0.98, len(k)) * (g[k + 1] - g[k])
    x = jit(gx, c[:, 0]).astype(np.float32)
    y = jit(gy, c[:, 1]).astype(np.float32)
    z = jit(gz, c[:, 2]).astype(np.float32)
    m = len(x)
    mtype = rng.integers(2, 7, m).astype(np.int8)
    mtype[::101] = 60
    x[::211] = -1.0
    markk = rng.uniform(300, 1600, m).astype(np.float32)
    markk[::53] = 0.0
    zdeep, tdeep = gy[-1] * 0.7, 1500.0
    for k, v in dict(HEATMIN=1e-2, t0koef=1.0, t1koef=1.0, heatdif=1.0, timestep=dt, maxtkstep=maxtkstep,
                     timesum=0.0, timebeg=1e30, timeend=2e30, tempend=0.0, zdeep=zdeep, tdeep=tdeep).items():
        R.set(k, v)
    for k, v in dict(multinumt=0, multicyct=1, frictyn=0, adiabyn=0).items():
        R.set(k, v, C.c_int)
    R.i32("multittt", 100)[0] = 2000  # the double iterate converges before the first check
    R.set("marknum", m, C.c_long)
    R.arr("markx", C.c_float, R.MAXMRK)[:m + 1] = np.r_[x, -1.0]
    R.arr("marky", C.c_float, R.MAXMRK)[:m + 1] = np.r_[y, -1.0]
    R.arr("markz", C.c_float, R.MAXMRK)[:m + 1] = np.r_[z, -1.0]
    R.arr("markt", C.c_byte, R.MAXMRK)[:m + 1] = np.r_[mtype, 2]
    R.arr("markk", C.c_float, R.MAXMRK)[:m + 1] = np.r_[markk, 0.0]
    for name, a in dict(tk=tk, ro=ro, cp=cp, kt=kt, ht=ht).items():
        R.f32(name)[:N] = a
    R.f32("et")[:N] = 0
    R.f64("pr")[:N] = 0
    R.lib.titerate(C.c_int(0))
    tk_out = R.f32("tk")[:N].copy()
    markk_out = R.arr("markk", C.c_float, R.MAXMRK)[:m].copy()
    np.savez_compressed(os.path.join(HERE, "g6_heat.npz"), n=n, gx=gx, gy=gy, gz=gz, bondm=cfg["bondm"],
                        bondv=cfg["bondv"], bondn=cfg["bondn"], ro=ro, cp=cp, kt=kt, ht=ht, tk=tk, dt=dt,
                        maxtkstep=maxtkstep, heatmin=1e-2, sweeps=2000, x=x, y=y, z=z, mtype=mtype, markk=markk, zdeep=zdeep,
                        tdeep=tdeep, heatdif=1.0, tk_out=tk_out, markk_out=markk_out)
    print("g6_heat.npz written")


def g7(R):
    """Value-mode strain rates/stresses (move3mg.c:3-710) and the
    residual-mode rows behind vpiterate's native norms (continerr,
    x/y/zstokserr with ynerr=1, bonderr; move3mg.c:718-1579, 2405-2560) on a
    random non-uniform grid with random fields."""
    import ctypes as C
    rng = np.random.default_rng(707)
    nx, ny, nz = 12, 10, 11
    gx = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, nx - 1)]) * 3e3
    gy = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, ny - 1)]) * 3e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.6, 1.4, nz - 1)]) * 3e3
    N = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, 0)
    rules, nr = I.freeslip_rules()
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    f = {k: (10 ** rng.uniform(19, 23, N)).astype(np.float32) for k in ("nu", "nxx", "nxy", "nxz", "nyz")}
    f["ro"] = rng.uniform(3000, 3400, N).astype(np.float32)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(bondm, bondv, bondn)
    R.set_physics(g=(0.3, 9.81, -0.2), nubeg=1e19, nuend=1e23)
    for k, v in f.items():
        R.f32(k)[:N] = v
    vel = {k: (rng.standard_normal(N) * 1e-9).astype(np.float32) for k in ("vx", "vy", "vz")}
    pr = rng.standard_normal(N) * 1e8
    for k, v in vel.items():
        R.f32(k)[:N] = v
    R.f64("pr")[:N] = pr
    sol = np.stack([pr, vel["vx"].astype(np.float64), vel["vy"].astype(np.float64),
                    vel["vz"].astype(np.float64)], 1).ravel()
    R.f64("sol0", R.MAXPPP)[:4 * N] = sol
    R.lib.refh_fill_stress()
    names = ["exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz"]
    for k in names:
        R.f32(k, R.MAXNNN)[:N] = 0
    R.lib.refh_fill_stress()
    stress = np.stack([R.f32(k, R.MAXNNN)[:N].copy() for k in names])
    rows = np.full((5, 4 * N, 2), np.nan)
    out = (C.c_double * 2)()
    for m3 in range(nz):
        for m1 in range(nx):
            for m2 in range(ny):
                n = m3 * nx * ny + m1 * ny + m2
                if m1 and m2 and m3:
                    R.lib.refh_err_row(0, m1, m2, m3, out)
                    rows[0, 4 * n] = list(out)
                for c, ok in ((1, m2 < ny - 1 and m3 < nz - 1), (2, m1 < nx - 1 and m3 < nz - 1),
                              (3, m1 < nx - 1 and m2 < ny - 1)):
                    if not ok:
                        continue
                    if bondm[4 * n + c] == 0:
                        R.lib.refh_err_row(c, m1, m2, m3, out)
                        rows[c, 4 * n + c] = list(out)
                    else:
                        R.lib.refh_err_row(4, 4 * n + c, 0, 0, out)
                        rows[4, 4 * n + c] = list(out)
    np.savez_compressed(os.path.join(HERE, "g7_stress_norms.npz"), gx=gx, gy=gy, gz=gz, bondm=bondm, bondv=bondv,
                        bondn=bondn, g=np.array([0.3, 9.81, -0.2]), nubeg=1e19, nuend=1e23, **f, **vel, pr=pr,
                        stress=stress, rows=rows)
    print("g7_stress_norms.npz written")


def g8(R, n=17):
    """heaterr's right-hand side with adiabatic and shear heating
    (heat3mg.c:54-160) at every interior node for (adiabyn, frictyn) =
    (1, 0), (2, 1), (0, 1), and one titerate step (heat3mg.c:287-588) with
    adiabyn 2 and frictyn 1, through the reference on configuration-0
    geometry with random fields (some rock below 2000 kg/m^3, a top layer
    above ysize - stp100)."""
    import ctypes as C
    rng = np.random.default_rng(808)
    cfg = I.make_config(0, n, n, n, seed=1)
    gx, gy, gz = cfg["gx"], cfg["gy"], cfg["gz"]
    N = n ** 3
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(cfg["bondm"], cfg["bondv"], cfg["bondn"])
    f32 = lambda a: np.asarray(a, np.float32)
    ro = f32(rng.uniform(1800, 3400, N))
    cp = f32(rng.uniform(900, 1300, N))
    kt = f32(rng.uniform(1.5, 4.5, N))
    ht = f32(rng.uniform(0, 2e-6, N))
    tk = f32(cfg["tk"] + rng.uniform(-50, 50, N))
    et = f32(rng.uniform(1e-5, 4e-5, N))
    vel = {k: f32(rng.standard_normal(N) * 1e-9) for k in ("vx", "vy", "vz")}
    yy = np.tile(gy, n * n)  # y of node m3*n*n + m1*n + m2
    pr = 3300.0 * 9.81 * yy + rng.uniform(-1e6, 1e6, N)
    s6 = {k: f32(rng.standard_normal(N) * 1e6) for k in ("sxx", "syy", "szz", "sxy", "sxz", "syz")}
    e6 = {k: f32(rng.standard_normal(N) * 1e-14) for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")}
    gk = np.array([0.3, 9.81, -0.2])
    ysize, stp100 = float(gy[-1]), float(0.15 * gy[-1])
    for k, v in dict(ysize=ysize, stp100=stp100, GXKOEF=gk[0], GYKOEF=gk[1], GZKOEF=gk[2]).items():
        R.set(k, v)
    for name, a in dict(tk=tk, tk0=tk, ro=ro, cp=cp, kt=kt, ht=ht, et=et, **vel, **s6, **e6).items():
        R.f32(name)[:N] = a
    R.f64("pr")[:N] = pr
    bt = cfg["bondm"][8 * N:9 * N]
    out = (C.c_double * 2)()
    res = {}
    for a, fr in ((1, 0), (2, 1), (0, 1)):
        R.set("adiabyn", a, C.c_int)
        R.set("frictyn", fr, C.c_int)
        rate = np.zeros(N)
        src = np.zeros(N, np.float32)
        for m3 in range(n):
            for m1 in range(n):
                for m2 in range(n):
                    v = m3 * n * n + m1 * n + m2
                    if bt[v] != 0:
                        continue
                    R.lib.refh_heat_rhs(C.c_long(m1), C.c_long(m2), C.c_long(m3), out)
                    rate[v], src[v] = out[0], out[1]
        res[f"rate_{a}{fr}"] = rate
        res[f"src_{a}{fr}"] = src
    # one implicit step with the terms on, three sub-steps
    dt = 1e13
    maxtkstep = np.abs(res["rate_21"]).max() * dt / 2.5
    for k, v in dict(HEATMIN=1e-2, t0koef=1.0, t1koef=1.0, heatdif=0.0, timestep=dt, maxtkstep=maxtkstep,
                     timesum=0.0, timebeg=1e30, timeend=2e30, tempend=0.0).items():
        R.set(k, v)
    for k, v in dict(multinumt=0, multicyct=1, frictyn=1, adiabyn=2).items():
        R.set(k, v, C.c_int)
    R.i32("multittt", 100)[0] = 2000
    R.set("marknum", 0, C.c_long)
    R.f32("tk")[:N] = tk
    R.lib.titerate(C.c_int(0))
    tk_out = R.f32("tk")[:N].copy()
    np.savez_compressed(os.path.join(HERE, "g8_heat_sources.npz"), n=n, gx=gx, gy=gy, gz=gz, bondm=cfg["bondm"],
                        bondv=cfg["bondv"], bondn=cfg["bondn"], ro=ro, cp=cp, kt=kt, ht=ht, tk=tk, et=et, pr=pr,
                        gk=gk, ysize=ysize, stp100=stp100, dt=dt, maxtkstep=maxtkstep, heatmin=1e-2, sweeps=2000,
                        tk_out=tk_out, **vel, **s6, **e6, **res)
    print("g8_heat_sources.npz written")


def g9(R):
    """viscmod 1 (geometric) and 2 (harmonic) averaging through the
    reference: the level-0 and level-1 Stokes rows (sxxcalc..syzcalc
    viscosity means, move3mg.c:36-48, 402-414, 517-529, 632-644) and
    ronurestrict (move3mg.c:2642-2741) on the g1 grid and fields, and
    ronurecalc's node fields (mark3mg.c:848-981) on the g5 markers."""
    import ctypes as C
    g = np.load(os.path.join(HERE, "g1_operator.npz"))
    rng = np.random.default_rng(909)
    nx, ny, nz, mn = int(g["nx"]), int(g["ny"]), int(g["nz"]), int(g["multinum"])
    N0 = nx * ny * nz
    R.setup_grid(g["gx"], g["gy"], g["gz"], mn)
    R.set_bc(g["bondm"], g["bondv"], g["bondn"])
    out = {}
    for vm in (1, 2):
        R.set_physics(g=(0.3, 9.81, -0.2), nubeg=1e19, nuend=1e23, viscmod=vm)
        for k in ("nu", "nxx", "nxy", "nxz", "nyz", "ro"):
            R.f32(k)[:N0] = g[k]
        A0, rhs0, ex0 = reference_system(R, 0)
        x0 = rng.standard_normal(4 * N0) * np.tile([1e7, 1e-9, 1e-9, 1e-9], N0)
        x0[~ex0] = 0
        R.lib.refh_ronurestrict(1)
        p1, N1 = int(R.mgp[1]), int(R.mgn[1])
        nu1 = R.f32("nu")[p1:p1 + N1].copy()
        A1, rhs1, ex1 = reference_system(R, 1)
        x1 = rng.standard_normal(4 * N1) * np.tile([1e7, 1e-9, 1e-9, 1e-9], N1)
        x1[~ex1] = 0
        b1 = rng.standard_normal(4 * N1)
        b1[~ex1] = 0
        out.update({f"x0_{vm}": x0, f"rhs0_{vm}": rhs0, f"r0_{vm}": rhs0 - A0 @ x0, f"exists0_{vm}": ex0,
                    f"nu1_{vm}": nu1, f"x1_{vm}": x1, f"b1_{vm}": b1, f"r1_{vm}": b1 - A1 @ x1,
                    f"exists1_{vm}": ex1})
    # ronurecalc with the g5 markers
    m5 = np.load(os.path.join(HERE, "g5_markers.npz"))
    gx, gy, gz = m5["gx"], m5["gy"], m5["gz"]
    N = len(gx) * len(gy) * len(gz)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(m5["bondm"], m5["bondv"], m5["bondn"])
    n = len(m5["x"])
    gnames = ["nu", "ro", "cp", "kt", "ht", "et", "tk", "nxx", "nxy", "nxz", "nyz"]
    names = dict(nu0="marknu", dh="markdh", dv="markdv", n0="markn0", n1="markn1", ro="markro", bro="markbb",
                 aro="markaa", cp="markcp", kt="markkt", kf="markkf", kp="markkp", ht="markht")
    fields = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")
    for vm in (1, 2):
        R.set("marknum", n, C.c_long)
        R.arr("markx", C.c_float, R.MAXMRK)[:n] = m5["x"]
        R.arr("marky", C.c_float, R.MAXMRK)[:n] = m5["y"]
        R.arr("markz", C.c_float, R.MAXMRK)[:n] = m5["z"]
        R.arr("markk", C.c_float, R.MAXMRK)[:n] = m5["tk"]
        R.arr("markv", C.c_float, R.MAXMRK)[:n] = m5["eta"]
        R.arr("markt", C.c_byte, R.MAXMRK)[:n] = m5["mtype"]
        R.arr("markrr", C.c_float, R.MAXMRK)[:n] = 0
        R.arr("marke", C.c_float, R.MAXMRK)[:n] = 0
        for q, f in enumerate(fields):
            a = R.arr(names[f], C.c_double, R.MAXTMR)
            a[:] = 0
            a[:len(m5["rocks"])] = m5["rocks"][:, q]
        for gname in ("markss", "markll", "marka0", "marka1", "markb0", "markb1", "marke0", "marke1"):
            R.arr(gname, C.c_double, R.MAXTMR)[:] = 0
        R.arr("markmm", C.c_double, R.MAXTMR)[:] = 1.0
        R.set("densimod", 1, C.c_int)
        for k, v in dict(eroslev=-1e30, sedilev=1e30, waterlev=1e30, vdeep=float(m5["vdeep"]),
                         nudeep=float(m5["nudeep"]), nukoef=0.0, nubeg=float(m5["nubeg"]), nuend=float(m5["nuend"]),
                         timestep=float(m5["timestep"])).items():
            R.set(k, v)
        R.set("viscmod", vm, C.c_int)
        R.f64("pr")[:N] = m5["pr"]
        for q, gname in enumerate(gnames):
            R.f32(gname)[:N] = m5["init"][q]
        R.lib.refh_ronurecalc()
        out[f"mk_out_{vm}"] = np.stack([R.f32(gname)[:N].copy() for gname in gnames])
    R.set("viscmod", 0, C.c_int)
    np.savez_compressed(os.path.join(HERE, "g9_viscmod.npz"), **out)
    print("g9_viscmod.npz written")


def g10(R):
    """Run with stdin closed (< /dev/null): with random, unrelated stress
    and strain-rate fields the dissipation can be negative, where epscalc1
    prints and waits for a key (mark3mg.c:2834).

    epscalc1 (mark3mg.c:2535-2838) at random points of the g5 grid,
    including points outside the grid and on grid and mid lines, and one
    movemark step (mark3mg.c:2-397) of the g5 markers: viscosity from
    viscalc at the marker pressure of epscalc1, accumulated strain with
    healing, RK4 positions."""
    import ctypes as C
    m5 = np.load(os.path.join(HERE, "g5_markers.npz"))
    rng = np.random.default_rng(1010)
    gx, gy, gz = m5["gx"], m5["gy"], m5["gz"]
    nx, ny, nz = len(gx), len(gy), len(gz)
    N = nx * ny * nz
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(m5["bondm"], m5["bondv"], m5["bondn"])
    names = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    st = {k: (rng.standard_normal(N) * (1e-14 if k[0] == "e" else 1e6)).astype(np.float32) for k in names}
    pr = rng.uniform(1e7, 1e9, N)
    for k, a in st.items():
        R.f32(k)[:N] = a
    R.f64("pr")[:N] = pr
    npt = 600
    pts = np.stack([rng.uniform(-2e3, gx[-1] + 2e3, npt), rng.uniform(-2e3, gy[-1] + 2e3, npt),
                    rng.uniform(-2e3, gz[-1] + 2e3, npt)], 1)
    pts[:30, 0] = gx[rng.integers(0, nx, 30)]
    pts[30:60, 1] = (gy[:-1] + gy[1:])[rng.integers(0, ny - 1, 30)] / 2.0
    pts[60:90, 2] = gz[rng.integers(0, nz, 30)]
    buf = (C.c_double * 14)()
    eps = np.zeros((npt, 14))
    for q, p in enumerate(pts):
        R.lib.refh_epscalc1(C.c_double(p[0]), C.c_double(p[1]), C.c_double(p[2]), buf)
        eps[q] = list(buf)
    # one movemark step of the g5 markers
    n = len(m5["x"])
    marke = np.where(rng.uniform(size=n) < 0.4, 0.0, rng.uniform(-0.3, 0.6, n)).astype(np.float32)
    markv = m5["eta"].copy()
    vel = {k: (rng.standard_normal(N) * 1e-9).astype(np.float32) for k in ("vx", "vy", "vz")}
    for k, a in vel.items():
        R.f32(k)[:N] = a
    R.set("marknum", n, C.c_long)
    R.arr("markx", C.c_float, R.MAXMRK)[:n] = m5["x"]
    R.arr("marky", C.c_float, R.MAXMRK)[:n] = m5["y"]
    R.arr("markz", C.c_float, R.MAXMRK)[:n] = m5["z"]
    R.arr("markk", C.c_float, R.MAXMRK)[:n] = m5["tk"]
    R.arr("markv", C.c_float, R.MAXMRK)[:n] = markv
    R.arr("markt", C.c_byte, R.MAXMRK)[:n] = m5["mtype"]
    R.arr("marke", C.c_float, R.MAXMRK)[:n] = marke
    for gname in ("markd", "markw", "markex", "markft", "markrr", "markto"):
        R.arr(gname, C.c_float, R.MAXMRK)[:n] = 0
    rnames = dict(nu0="marknu", dh="markdh", dv="markdv", n0="markn0", n1="markn1", ro="markro", bro="markbb",
                  aro="markaa", cp="markcp", kt="markkt", kf="markkf", kp="markkp", ht="markht")
    fields = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")
    for q, f in enumerate(fields):
        a = R.arr(rnames[f], C.c_double, R.MAXTMR)
        a[:] = 0
        a[:len(m5["rocks"])] = m5["rocks"][:, q]
    for gname in ("markss", "markll", "marka0", "marka1", "markb0", "markb1", "marke0", "marke1"):
        R.arr(gname, C.c_double, R.MAXTMR)[:] = 0
    R.arr("markmm", C.c_double, R.MAXTMR)[:] = 1.0
    dt = 3e11
    for k, v in dict(timestep=dt, healrate=2e-15, nubeg=float(m5["nubeg"]), nuend=float(m5["nuend"]), zdeep=1e30,
                     vdeep=1e30, tdeep=0.0, stp100=1e30, vyfluid=0.0, vymelt=0.0,
                     timesum=0.0).items():
        R.set(k, v)
    R.set("outgrid", 0, C.c_int)
    R.set("markmod", 0, C.c_int)
    R.lib.refh_movemark()
    out = dict(x_out=R.arr("markx", C.c_float, R.MAXMRK)[:n].copy(), y_out=R.arr("marky", C.c_float, R.MAXMRK)[:n].copy(),
               z_out=R.arr("markz", C.c_float, R.MAXMRK)[:n].copy(), tk_out=R.arr("markk", C.c_float, R.MAXMRK)[:n].copy(),
               eta_out=R.arr("markv", C.c_float, R.MAXMRK)[:n].copy(), e_out=R.arr("marke", C.c_float, R.MAXMRK)[:n].copy(),
               type_out=R.arr("markt", C.c_byte, R.MAXMRK)[:n].copy())
    np.savez_compressed(os.path.join(HERE, "g10_eps_movemark.npz"), pts=pts, eps=eps, pr=pr, **st, **vel,
                        marke=marke, markv=markv, dt=dt, healrate=2e-15, **out)
    print("g10_eps_movemark.npz written")


def g11(R):
    """One movemark step (mark3mg.c:2-397) of the g5 markers with the depth
    rules active (zdeep, vdeep inside the grid, tdeep above some marker
    temperatures) and velocities large enough for markers to cross them;
    inputs as g10 otherwise.  Run with stdin closed (see g10)."""
    import ctypes as C
    m5 = np.load(os.path.join(HERE, "g5_markers.npz"))
    r10 = np.load(os.path.join(HERE, "g10_eps_movemark.npz"))
    rng = np.random.default_rng(1111)
    gx, gy, gz = m5["gx"], m5["gy"], m5["gz"]
    N = len(gx) * len(gy) * len(gz)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(m5["bondm"], m5["bondv"], m5["bondn"])
    names = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    for k in names:
        R.f32(k)[:N] = r10[k]
    R.f64("pr")[:N] = r10["pr"]
    vel = {k: (rng.standard_normal(N) * 2e-8).astype(np.float32) for k in ("vx", "vy", "vz")}
    for k, a in vel.items():
        R.f32(k)[:N] = a
    n = len(m5["x"])
    mtype = m5["mtype"].copy()
    R.set("marknum", n, C.c_long)
    R.arr("markx", C.c_float, R.MAXMRK)[:n] = m5["x"]
    R.arr("marky", C.c_float, R.MAXMRK)[:n] = m5["y"]
    R.arr("markz", C.c_float, R.MAXMRK)[:n] = m5["z"]
    R.arr("markk", C.c_float, R.MAXMRK)[:n] = m5["tk"]
    R.arr("markv", C.c_float, R.MAXMRK)[:n] = r10["markv"]
    R.arr("markt", C.c_byte, R.MAXMRK)[:n] = mtype
    R.arr("marke", C.c_float, R.MAXMRK)[:n] = r10["marke"]
    for gname in ("markd", "markw", "markex", "markft", "markrr", "markto"):
        R.arr(gname, C.c_float, R.MAXMRK)[:n] = 0
    rnames = dict(nu0="marknu", dh="markdh", dv="markdv", n0="markn0", n1="markn1", ro="markro", bro="markbb",
                  aro="markaa", cp="markcp", kt="markkt", kf="markkf", kp="markkp", ht="markht")
    fields = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")
    for q, f in enumerate(fields):
        a = R.arr(rnames[f], C.c_double, R.MAXTMR)
        a[:] = 0
        a[:len(m5["rocks"])] = m5["rocks"][:, q]
    for gname in ("markss", "markll", "marka0", "marka1", "markb0", "markb1", "marke0", "marke1"):
        R.arr(gname, C.c_double, R.MAXTMR)[:] = 0
    R.arr("markmm", C.c_double, R.MAXTMR)[:] = 1.0
    dt = 3e11
    zdeep, vdeep, tdeep = 0.55 * gy[-1], 0.7 * gy[-1], 1200.0
    for k, v in dict(timestep=dt, healrate=2e-15, nubeg=float(m5["nubeg"]), nuend=float(m5["nuend"]), zdeep=zdeep,
                     vdeep=vdeep, tdeep=tdeep, stp100=1e30, vyfluid=0.0, vymelt=0.0, timesum=0.0).items():
        R.set(k, v)
    R.set("outgrid", 0, C.c_int)
    R.set("markmod", 0, C.c_int)
    R.lib.refh_movemark()
    out = dict(x_out=R.arr("markx", C.c_float, R.MAXMRK)[:n].copy(), y_out=R.arr("marky", C.c_float, R.MAXMRK)[:n].copy(),
               z_out=R.arr("markz", C.c_float, R.MAXMRK)[:n].copy(), tk_out=R.arr("markk", C.c_float, R.MAXMRK)[:n].copy(),
               eta_out=R.arr("markv", C.c_float, R.MAXMRK)[:n].copy(), e_out=R.arr("marke", C.c_float, R.MAXMRK)[:n].copy(),
               type_out=R.arr("markt", C.c_byte, R.MAXMRK)[:n].copy())
    print("types changed:", int((out["type_out"] != mtype).sum()), "tk raised:",
          int((out["tk_out"] > m5["tk"]).sum()))
    np.savez_compressed(os.path.join(HERE, "g11_movemark_depth.npz"), **vel, dt=dt, healrate=2e-15, zdeep=zdeep,
                        vdeep=vdeep, tdeep=tdeep, **out)
    print("g11_movemark_depth.npz written")


# rock table fields in the order in3mg.c:204-231 reads them from a deck
DECK_ROCK_ORDER = ("n0", "n1", "s0", "s1", "nu0", "dh", "dv", "ss", "mm", "ll", "a0", "a1", "b0", "b1", "e0",
                   "e1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")
REF_ROCK_NAMES = dict(n0="markn0", n1="markn1", s0="marks0", s1="marks1", nu0="marknu", dh="markdh", dv="markdv",
                      ss="markss", mm="markmm", ll="markll", a0="marka0", a1="marka1", b0="markb0", b1="markb1",
                      e0="marke0", e1="marke1", ro="markro", bro="markbb", aro="markaa", cp="markcp", kt="markkt",
                      kf="markkf", kp="markkp", ht="markht")


def deck_rocks(path="/root/reference/init.t3c"):
    """The /ROCKS_DESCRIPTIONS/ table of the reference's deck as a list of
    dicts indexed by rock number: the numeric rows in the column order of
    in3mg.c:204-231 (whitespace tokens, '/' tokens and the column header
    skipped, up to '~')."""
    toks = [t for t in open(path).read().split("/ROCKS_DESCRIPTIONS/", 1)[1].split() if not t.startswith("/")]
    rocks = {}
    i = 0
    while not toks[i].isdigit():  # the column header line
        i += 1
    while toks[i] != "~":
        num = int(toks[i])
        vals = toks[i + 1:i + 1 + len(DECK_ROCK_ORDER)]
        rocks[num] = {f: float(v[1:]) if v.startswith("k") else float(v) for f, v in zip(DECK_ROCK_ORDER, vals)}
        i += 1 + len(DECK_ROCK_ORDER)
    out = [dict() for _ in range(max(rocks) + 1)]
    for k, r in rocks.items():
        out[k] = r
    return out


def _set_ref_rocks(R, rocks):
    import ctypes as C
    for f, gname in REF_ROCK_NAMES.items():
        a = R.arr(gname, C.c_double, R.MAXTMR)
        a[:] = 0
        for t, r in enumerate(rocks):
            a[t] = r.get(f, 0.0)


def _rock_matrix(rocks):
    import oracle
    return np.array([[r.get(f, 0.0) for f in oracle.ROCKX_FIELDS] for r in rocks])


def g14(R):
    """movemark (mark3mg.c:2-397) with the reference deck's rock table
    (init.t3c: power law, brittle strength, Karato & Wu, the grain-size
    model of the mantle rocks): marker viscosity, strain, grain roughness
    and dislocation share, positions, for (a) outgrid 2, RK4, timesum past
    convend and (b) outgrid 1, markmod 1, timesum 0; fluid markers (types
    50-99, with markd) and immobile markers (>= 100) included.  Run with
    stdin closed (see g10)."""
    import ctypes as C
    rng = np.random.default_rng(1414)
    rocks = deck_rocks()
    nx, ny, nz = 11, 9, 10
    gx = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nx - 1)]) * 5e3
    gy = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, ny - 1)]) * 5e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nz - 1)]) * 5e3
    N = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, 0)
    rules, nr = I.freeslip_rules(273.0, 1600.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    n = 1500
    x = rng.uniform(-3e3, gx[-1] + 3e3, n).astype(np.float32)
    y = rng.uniform(-3e3, gy[-1] + 3e3, n).astype(np.float32)
    z = rng.uniform(-3e3, gz[-1] + 3e3, n).astype(np.float32)
    solid = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 16, 17, 18])
    mtype = solid[rng.integers(0, len(solid), n)].astype(np.int8)
    mtype[rng.uniform(size=n) < 0.08] = rng.choice([52, 61, 75, 99], 1)[0]
    fl = rng.uniform(size=n) < 0.08
    mtype[fl] = rng.choice(np.array([52, 61, 75, 99], np.int8), fl.sum())
    im = rng.uniform(size=n) < 0.05
    mtype[im] = rng.choice(np.array([102, 109, 114, 127], np.int8), im.sum())
    tk = rng.uniform(400, 1900, n).astype(np.float32)
    tk[rng.uniform(size=n) < 0.04] = 0
    tk[fl & (rng.uniform(size=n) < 0.3)] = 250.0  # cold fluid: -|vy_water|
    eta = np.where(rng.uniform(size=n) < 0.3, 0.0, 10 ** rng.uniform(19, 24, n)).astype(np.float32)
    marke = np.where(rng.uniform(size=n) < 0.3, 0.0, rng.uniform(-0.8, 1.5, n)).astype(np.float32)
    markrr = np.where(rng.uniform(size=n) < 0.3, 0.0, 10 ** rng.uniform(-5.5, -2.5, n)).astype(np.float32)
    markrr[rng.uniform(size=n) < 0.05] = -1e-4
    markrn = rng.uniform(0, 1, n).astype(np.float32)
    markd = np.where(rng.uniform(size=n) < 0.5, 1000.0, 1250.0).astype(np.float32)
    names = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    st = {k: (rng.standard_normal(N) * (10 ** rng.uniform(-15.5, -13, N) if k[0] == "e" else 10 ** rng.uniform(6, 8, N))
              ).astype(np.float32) for k in names}
    pr = rng.uniform(1e7, 3e9, N)
    vel = {k: (rng.standard_normal(N) * 3e-9).astype(np.float32) for k in ("vx", "vy", "vz")}
    ro = rng.uniform(1000, 3300, N).astype(np.float32)
    cases = [dict(outgrid=2, markmod=0, timesum=5e13), dict(outgrid=1, markmod=1, timesum=0.0)]
    prm = dict(timestep=3e11, healrate=1e-14, nubeg=1e18, nuend=1e25, convend=1e13, vyfluid=1e-9, vymelt=3e-9,
               g=(0.0, 9.81, 0.0))
    outs = {}
    for ci, cs in enumerate(cases):
        R.setup_grid(gx, gy, gz, 0)
        R.set_bc(bondm, bondv, bondn)
        for k, a in st.items():
            R.f32(k)[:N] = a
        for k, a in vel.items():
            R.f32(k)[:N] = a
        R.f32("ro")[:N] = ro
        R.f64("pr")[:N] = pr
        R.arr("bufv", C.c_double, R.MAXPPP)[:] = 0
        R.set("marknum", n, C.c_long)
        R.arr("markx", C.c_float, R.MAXMRK)[:n] = x
        R.arr("marky", C.c_float, R.MAXMRK)[:n] = y
        R.arr("markz", C.c_float, R.MAXMRK)[:n] = z
        R.arr("markk", C.c_float, R.MAXMRK)[:n] = tk
        R.arr("markv", C.c_float, R.MAXMRK)[:n] = eta
        R.arr("markt", C.c_byte, R.MAXMRK)[:n] = mtype
        R.arr("marke", C.c_float, R.MAXMRK)[:n] = marke
        R.arr("markrr", C.c_float, R.MAXMRK)[:n] = markrr
        R.arr("markrn", C.c_float, R.MAXMRK)[:n] = markrn
        R.arr("markd", C.c_float, R.MAXMRK)[:n] = markd
        for gname in ("markw", "markex", "markft", "markto"):
            R.arr(gname, C.c_float, R.MAXMRK)[:n] = 0
        _set_ref_rocks(R, rocks)
        for k, v in dict(timestep=prm["timestep"], healrate=prm["healrate"], nubeg=prm["nubeg"], nuend=prm["nuend"],
                         zdeep=1e30, vdeep=1e30, tdeep=0.0, stp100=1e30, vyfluid=prm["vyfluid"],
                         vymelt=prm["vymelt"], timesum=cs["timesum"], convend=prm["convend"],
                         GXKOEF=prm["g"][0], GYKOEF=prm["g"][1], GZKOEF=prm["g"][2]).items():
            R.set(k, v)
        R.set("outgrid", cs["outgrid"], C.c_int)
        R.set("markmod", cs["markmod"], C.c_int)
        R.lib.refh_movemark()
        for k, gname, ct in (("x", "markx", C.c_float), ("y", "marky", C.c_float), ("z", "markz", C.c_float),
                             ("tk", "markk", C.c_float), ("eta", "markv", C.c_float), ("e", "marke", C.c_float),
                             ("rr", "markrr", C.c_float), ("rn", "markrn", C.c_float), ("type", "markt", C.c_byte)):
            outs[f"{k}_out{ci}"] = R.arr(gname, ct, R.MAXMRK)[:n].copy()
        assert R.get("marknum", C.c_long) == n
    np.savez_compressed(os.path.join(HERE, "g14_movemark_rheology.npz"), gx=gx, gy=gy, gz=gz, x=x, y=y, z=z,
                        mtype=mtype, tk=tk, eta=eta, marke=marke, markrr=markrr, markrn=markrn, markd=markd, pr=pr,
                        ro=ro, **st, **vel, rocks=_rock_matrix(rocks),
                        outgrid=np.array([c["outgrid"] for c in cases]),
                        markmod=np.array([c["markmod"] for c in cases]),
                        timesum=np.array([c["timesum"] for c in cases]), **{k: v for k, v in prm.items() if k != "g"},
                        g=np.array(prm["g"]), **outs)
    print("g14_movemark_rheology.npz written")


def g15(R):
    """ronurecalc (mark3mg.c:404-1015) with the reference deck's rocks:
    erosion/sedimentation and water/air retyping, markers across the water
    level (uniform y spacing), viscalc from epscalc1 strain rates
    (timestep 0), eclogite and mantle phase-transition densities with the
    teclmin/teclmax damping; densimod 1, no melt types."""
    import ctypes as C
    rng = np.random.default_rng(1515)
    rocks = deck_rocks()
    nx, ny, nz = 11, 9, 10
    gx = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nx - 1)]) * 5e3
    gy = np.arange(ny) * 5e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nz - 1)]) * 5e3
    N = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, 0)
    rules, nr = I.freeslip_rules(273.0, 1600.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(bondm, bondv, bondn)
    n = 2400
    x = rng.uniform(1.0, gx[-1] - 1.0, n).astype(np.float32)
    y = rng.uniform(1.0, gy[-1] - 1.0, n).astype(np.float32)
    z = rng.uniform(1.0, gz[-1] - 1.0, n).astype(np.float32)
    waterlev, eroslev, sedilev, mnumy = 11.3e3, 4.0e3, 9.0e3, 2
    near = rng.uniform(size=n) < 0.15
    y[near] = (waterlev + rng.uniform(-1.2, 1.2, near.sum()) * 1.25e3).astype(np.float32)
    solid = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 16, 17, 18])
    mtype = solid[rng.integers(0, len(solid), n)].astype(np.int8)
    mtype[near] = rng.choice(np.array([0, 1], np.int8), near.sum())
    tk = rng.uniform(500, 2000, n).astype(np.float32)
    tk[::113] = 0
    eta = np.where(rng.uniform(size=n) < 0.5, 0.0, 10 ** rng.uniform(19, 24, n)).astype(np.float32)
    marke = np.where(rng.uniform(size=n) < 0.3, 0.0, rng.uniform(-0.8, 1.5, n)).astype(np.float32)
    markrr = np.where(rng.uniform(size=n) < 0.3, 0.0, 10 ** rng.uniform(-5.5, -2.5, n)).astype(np.float32)
    markrn = rng.uniform(0, 1, n).astype(np.float32)
    names = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    st = {k: (rng.standard_normal(N) * (10 ** rng.uniform(-15.5, -13, N) if k[0] == "e" else 10 ** rng.uniform(6, 8, N))
              ).astype(np.float32) for k in names}
    pr = rng.uniform(1e8, 4e10, N)  # up to 400 kbar: every transition is crossed
    for k, a in st.items():
        R.f32(k)[:N] = a
    R.f64("pr")[:N] = pr
    R.set("marknum", n, C.c_long)
    for gname, a in (("markx", x), ("marky", y), ("markz", z), ("markk", tk), ("markv", eta), ("marke", marke),
                     ("markrr", markrr), ("markrn", markrn)):
        R.arr(gname, C.c_float, R.MAXMRK)[:n] = a
    R.arr("markt", C.c_byte, R.MAXMRK)[:n] = mtype
    for gname in ("markd", "markw", "markex", "markft", "markto"):
        R.arr(gname, C.c_float, R.MAXMRK)[:n] = 0
    _set_ref_rocks(R, rocks)
    R.set("densimod", 1, C.c_int)
    R.set("viscmod", 0, C.c_int)
    sedimnum, sedimcyc = 4, 3
    R.set("sedimnum", sedimnum, C.c_int)
    R.set("sedimcyc", sedimcyc, C.c_int)
    R.set("mnumy", mnumy, C.c_long)
    prm = dict(eroslev=eroslev, sedilev=sedilev, waterlev=waterlev, vdeep=gy[-1] * 0.8, nudeep=1e21, nukoef=0.0,
               nubeg=1e18, nuend=1e25, timestep=0.0, timesum=5e13, convend=1e13, teclmin=700.0, teclmax=1100.0)
    for k, v in prm.items():
        R.set(k, v)
    init = np.zeros((11, N), np.float32)
    init[0] = 5e20; init[1] = 3111; init[2] = 999; init[3] = 2.2; init[4] = 1e-9; init[5] = 1e-5
    init[6] = 1000; init[7:] = 7e20
    gnames = ["nu", "ro", "cp", "kt", "ht", "et", "tk", "nxx", "nxy", "nxz", "nyz"]
    for q, gname in enumerate(gnames):
        R.f32(gname)[:N] = init[q]
    R.lib.refh_ronurecalc()
    m1 = (sedimnum + 1) // sedimcyc
    sedtype = 3 if (m1 // 2) * 2 == m1 else 4
    out = np.stack([R.f32(gname)[:N].copy() for gname in gnames])
    res = {f"{k}_out": R.arr(g, C.c_float, R.MAXMRK)[:n].copy()
           for k, g in (("eta", "markv"), ("e", "marke"), ("rr", "markrr"), ("rn", "markrn"))}
    res["type_out"] = R.arr("markt", C.c_byte, R.MAXMRK)[:n].copy()
    np.savez_compressed(os.path.join(HERE, "g15_ronurecalc_rheology.npz"), gx=gx, gy=gy, gz=gz, bondm=bondm,
                        bondv=bondv, bondn=bondn, x=x, y=y, z=z,
                        mtype=mtype, tk=tk, eta=eta, marke=marke, markrr=markrr, markrn=markrn, pr=pr, **st,
                        rocks=_rock_matrix(rocks), init=init, out=out, sedtype=sedtype, mnumy=mnumy,
                        densimod=1, **prm, **res)
    print("g15_ronurecalc_rheology.npz written", "sedtype", sedtype)


def g17(R):
    """ronurecalc (mark3mg.c:404-1015) on partially molten rocks: the deck's
    melt types (23-30, 32, 34, 36-38) and their solid parents at
    temperatures and pressures across every family's solidus and liquidus,
    dry and wet (markw) markers, extracted melt (markex), the dry mantle
    pressure limit maxpmelt; densimod 0 and 1."""
    import ctypes as C
    rng = np.random.default_rng(1717)
    rocks = deck_rocks()
    nx, ny, nz = 9, 11, 8
    gx = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nx - 1)]) * 5e3
    gy = np.arange(ny) * 5e3
    gz = np.cumsum(np.r_[0, rng.uniform(0.7, 1.3, nz - 1)]) * 5e3
    N = nx * ny * nz
    h = I.hierarchy(nx, ny, nz, 0)
    rules, nr = I.freeslip_rules(273.0, 1600.0)
    bondm, bondv, bondn = I.expand_bc(rules, nr, h)
    R.setup_grid(gx, gy, gz, 0)
    R.set_bc(bondm, bondv, bondn)
    n = 3000
    x = rng.uniform(1.0, gx[-1] - 1.0, n).astype(np.float32)
    y = rng.uniform(1.0, gy[-1] - 1.0, n).astype(np.float32)
    z = rng.uniform(1.0, gz[-1] - 1.0, n).astype(np.float32)
    melt = np.array([23, 24, 25, 26, 27, 28, 29, 30, 32, 34, 36, 37, 38])
    solid = np.array([2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14])
    mtype = np.where(rng.uniform(size=n) < 0.8, melt[rng.integers(0, len(melt), n)],
                     solid[rng.integers(0, len(solid), n)]).astype(np.int8)
    tk = rng.uniform(800, 2200, n).astype(np.float32)
    eta = np.where(rng.uniform(size=n) < 0.5, 0.0, 10 ** rng.uniform(17, 23, n)).astype(np.float32)
    marke = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0, 1.5, n)).astype(np.float32)
    markrr = np.where(rng.uniform(size=n) < 0.3, 0.0, 10 ** rng.uniform(-5.5, -2.5, n)).astype(np.float32)
    markrn = rng.uniform(0, 1, n).astype(np.float32)
    markw = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0.01, 3.0, n)).astype(np.float32)
    markex = np.where(rng.uniform(size=n) < 0.6, 0.0, rng.uniform(0, 0.4, n)).astype(np.float32)
    names = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    st = {k: (rng.standard_normal(N) * (10 ** rng.uniform(-15.5, -13, N) if k[0] == "e" else 10 ** rng.uniform(6, 8, N))
              ).astype(np.float32) for k in names}
    # 0.1-60 kbar: melting intervals of every family, the dry mantle limit
    # (maxpmelt 3 GPa) and the eclogite/spinel transitions of densimod 1
    pr = 10 ** rng.uniform(7, np.log10(6e9), N)
    maxpmelt = 3e9
    out_all = {}
    for densimod in (0, 1):
        for k, a in st.items():
            R.f32(k)[:N] = a
        R.f64("pr")[:N] = pr
        R.set("marknum", n, C.c_long)
        for gname, a in (("markx", x), ("marky", y), ("markz", z), ("markk", tk), ("markv", eta),
                         ("marke", marke), ("markrr", markrr), ("markrn", markrn), ("markw", markw),
                         ("markex", markex)):
            R.arr(gname, C.c_float, R.MAXMRK)[:n] = a
        R.arr("markt", C.c_byte, R.MAXMRK)[:n] = mtype
        for gname in ("markd", "markft", "markto"):
            R.arr(gname, C.c_float, R.MAXMRK)[:n] = 0
        _set_ref_rocks(R, rocks)
        R.set("densimod", densimod, C.c_int)
        R.set("viscmod", 0, C.c_int)
        R.set("sedimnum", 4, C.c_int)
        R.set("sedimcyc", 3, C.c_int)
        R.set("mnumy", 2, C.c_long)
        prm = dict(eroslev=-1e30, sedilev=1e30, waterlev=-1.0, vdeep=1e30, nudeep=1e30, nukoef=0.0,
                   nubeg=1e17, nuend=1e25, timestep=0.0, timesum=5e13, convend=1e13, teclmin=700.0, teclmax=1100.0,
                   maxpmelt=maxpmelt)
        for k, v in prm.items():
            R.set(k, v)
        init = np.zeros((11, N), np.float32)
        gnames = ["nu", "ro", "cp", "kt", "ht", "et", "tk", "nxx", "nxy", "nxz", "nyz"]
        for q, gname in enumerate(gnames):
            R.f32(gname)[:N] = init[q]
        R.lib.refh_ronurecalc()
        out_all[f"out{densimod}"] = np.stack([R.f32(gname)[:N].copy() for gname in gnames])
        for k, g in (("eta", "markv"), ("e", "marke"), ("rr", "markrr"), ("rn", "markrn")):
            out_all[f"{k}_out{densimod}"] = R.arr(g, C.c_float, R.MAXMRK)[:n].copy()
        out_all[f"type_out{densimod}"] = R.arr("markt", C.c_byte, R.MAXMRK)[:n].copy()
    np.savez_compressed(os.path.join(HERE, "g17_ronurecalc_melt.npz"), gx=gx, gy=gy, gz=gz, bondm=bondm,
                        bondv=bondv, bondn=bondn, x=x, y=y, z=z, mtype=mtype, tk=tk, eta=eta, marke=marke,
                        markrr=markrr, markrn=markrn, markw=markw, markex=markex, pr=pr, **st,
                        rocks=_rock_matrix(rocks), sedtype=3, mnumy=2, **prm, **out_all)
    print("g17_ronurecalc_melt.npz written")


G16_ROCKS = {
    # type: n0 n1 s0 s1 nu dh dv ss mm ll a0 a1 b0 b1 e0 e1 ro bb aa cp kt kf kp ht (own values)
    0: "1e18 1e18 0 5e29 1e18 0 0 0 1 0 0 0 0 0 0 0.5 1 0 0 3.3e6 200 0 0 0",
    1: "1e18 1e18 0 5e29 1e18 0 0 0 1 0 0 0 0 0 0 0.5 1000 0 0 3.3e3 200 0 0 0",
    2: "1e18 1e26 2 5e29 1.9e17 1.5e5 0 3e4 2.4 1 2e6 2e6 0 0 0 0.5 2650 2e-5 0.5e-3 1000 0.65 800 0 2e-6",
    5: "1e18 1e26 0 5e29 1.9e17 1.5e5 0 3e4 2.4 1 3e6 3e6 0.3 0 0 0.5 2750 2e-5 0.6e-3 1000 0.65 800 0 2e-6",
    6: "1e18 1e26 0 5e29 4.7e22 2.4e5 0 3e4 3.1 1 3e6 3e6 0.3 0 0 0.5 2950 2e-5 0.6e-3 1000 1.2 470 0 2e-7",
    7: "1e18 1e26 0 5e29 4.7e22 2.4e5 0 3e4 3.1 1 3e6 3e6 0 0 0 0.5 3000 2e-5 0.6e-3 1000 1.2 470 0 2.2e-7",
    8: "1e18 1e26 0 5e29 4.7e22 2.4e5 0 3e4 3.1 1 3e6 3e6 0.6 0 0 0.5 3000 2e-5 0.6e-3 1000 1.2 470 0 2.4e-7",
    9: "1e18 1e26 0 1.4e-15 1.2e-16 -5.2e5 2.5 3.7e5 3.5 0.7 3e6 3e6 0.6 0 0 0.5 3300 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.2e-8",
    10: "1e18 1e26 0 1.4e-15 1.2e-16 -5.2e5 2.5 3.7e5 3.5 0.7 3e6 3e6 0.6 0 0 0.5 3300 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.4e-8",
    11: "1e18 1e26 0 5e15 -2e18 4.2e5 1 2.4e5 3 0.4 3e6 3e6 0 0 0 1 3250 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.2e-8",
    12: "1e18 1e26 0 1.4e-15 1.2e-16 -5.2e5 2.5 3.7e5 3.5 0.7 3e6 3e6 0 0 0 0.5 3300 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.6e-8",
    13: "1e18 1e26 0 5e29 3e36 9e3 0.3 3e6 3.8 1 3e6 3e6 0 0 0 1 3200 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.2e-8",
    14: "1e18 1e26 0 8e15 -3e22 5.3e5 2 3e5 3.5 0.4 3e6 3e6 0.6 0 0 0.5 3300 2e-5 0.6e-3 1000 0.73 1290 4e-6 2.2e-8",
}

# rock boxes: type, melt extraction, strain, 8 corners (x y z), fractions of
# the size or 'm'-prefixed metres (in3mg.c:550-575)
G16_RBOX = [
    "10 0.00 0.00  -0.1 -0.1 -0.1  -0.1 1.1 -0.1  1.1 -0.1 -0.1  1.1 1.1 -0.1"
    "  -0.1 -0.1 1.1  -0.1 1.1 1.1  1.1 -0.1 1.1  1.1 1.1 1.1",
    "9 0.05 0.10  0 m18000 0  0 m62000 0  1 m22000 0  1 m58000 0  0 m18000 1  0 m62000 1  1 m22000 1  1 m58000 1",
    "5 0.00 0.25  0 0 0  0 m18000 0  1 0 0  1 m22000 0  0 0 1  0 m18000 1  1 0 1  1 m22000 1",
    "2 0.00 0.00  0 0 0  0 m7000 0  1 0 0  1 m7000 0  0 0 1  0 m7000 1  1 0 1  1 m7000 1",
    "12 0.10 0.70  m90000 m18000 m40000  m130000 m80000 m40000  m110000 m18000 m40000  m150000 m80000 m40000"
    "  m70000 m18000 m140000  m115000 m80000 m140000  m95000 m18000 m140000  m135000 m80000 m140000",
    "7 0.00 0.00  m90000 m18000 m40000  m96000 m30000 m40000  m110000 m18000 m40000  m116000 m30000 m40000"
    "  m70000 m18000 m140000  m76000 m30000 m140000  m95000 m18000 m140000  m101000 m30000 m140000",
    "s11 0.20 0.05  0.62 0.70 0.45  0 0 0  0.0 0 0  0 0 0  0 0 0  0.09 0 0  0 0 0  1.0 1.2 0.9",
    "c13 0.00 0.30  0.25 0.80 0.60  0 0 0  0.03 0 0  0 0 0  0 0 0  0.11 0 0  0 0 0  1.0 0.15 1.0",
    "114 0.00 0.00  m160000 m70000 m20000  m160000 m100000 m20000  m200000 m70000 m20000  m200000 m100000 m20000"
    "  m160000 m70000 m60000  m160000 m100000 m60000  m200000 m70000 m60000  m200000 m100000 m60000",
    "8 0.00 0.00  0.05 0.40 0.10  0.05 0.47 0.10  0.20 0.40 0.12  0.20 0.47 0.12  0.05 0.40 0.30  0.05 0.47 0.30"
    "  0.20 0.40 0.33  0.20 0.47 0.33",
]

# temperature boxes (in3mg.c:806-951): type x0 x2 y0 y1 y2 y3 z0 x5 x7 y4 y5 y6 y7 z7 t0..t7
G16_TBOX = [
    "0  -0.000001 1.000001  -0.000001 1.000001 -0.000001 1.000001 -0.000001  -0.000001 1.000001 -0.000001 1.000001"
    " -0.000001 1.000001 1.000001  273 1700 273 1700 273 1700 273 1700",
    "4  -0.000001 0.55  m7000 m80000 m7000 m80000 -0.000001  -0.000001 0.50 m7000 m80000 m7000 m80000 1.000001"
    "  273 1580 3e7 5e7 3e7 5e7 1e-6 0.4e-3",
    "4  0.75 1.000001  m7000 m80000 m7000 m80000 -0.000001  0.78 1.000001 m7000 m80000 m7000 m80000 1.000001"
    "  273 1580 8e7 8e7 6e7 6e7 1e-6 0.4e-3",
    "5  0.55 0.75  m7000 m80000 m7000 m80000 -0.000001  0.50 0.78 m7000 m80000 m7000 m80000 1.000001"
    "  0 0 0 0 0 0 0 0",
    "2  0.62 m6000  0.70 0 0 0 0.45  m21000 1.0 0 0 0 1.2 0.9  1750 1400 0 0 0 0 0 0",
    "3  0.25 m3300  0.80 0 0 0 0.60  m12000 1.0 0 0 0 0.15 1.0  1300 900 0 0 0 0 0 0",
]

G16_BC = """ P     1   x     1   1     1   1     0       1.0 +0 +0 +1   0 0
 P     1   1     1   1     2   z-1   0       1.0 +1 +0 +0   0 0
 P     1   x     1   1     z   z     0       1.0 +0 +0 -1   0 0
 P     x   x     1   1     2   z-1   0       1.0 -1 +0 +0   0 0
 Vx    1   x-1   0   0     0   z-1   0       1.0 +0 +1 +0   0 0
 Vy    0   x-1   0   0     0   z-1   0       0   0 0
 Vz    0   x-1   0   0     1   z-1   0       1.0 +0 +1 +0   0 0
 T     0   x     0   0     0   z     273     0   0 0
 P     1   x     y   y     1   1     0       1.0 +0 +0 +1   0 0
 P     1   1     y   y     2   z-1   0       1.0 +1 +0 +0   0 0
 P     1   x     y   y     z   z     0       1.0 +0 +0 -1   0 0
 P     x   x     y   y     2   z-1   0       1.0 -1 +0 +0   0 0
 Vx    1   x-1   y-1 y-1   0   z-1   0       1.0 +0 -1 +0   0 0
 Vy    0   x-1   y   y     0   z-1   0       0   0 0
 Vz    0   x-1   y-1 y-1   1   z-1   0       1.0 +0 -1 +0   0 0
 T     0   x     y   y     0   z     0       1.0 +0 -1 +0   0 0
 P     1   1     2   y-1   1   1     0       1.0 +1 +0 +1   0 0
 P     1   1     2   y-1   z   z     0       1.0 +1 +0 -1   0 0
 Vx    0   0     0   y-1   0   z-1   0       0   0 0
 Vy    0   0     1   y-1   0   z-1   0       1.0 +1 +0 +0   0 0
 Vz    0   0     0   y-1   1   z-1   0       1.0 +1 +0 +0   0 0
 T     0   0     1   y-1   0   z     0       1.0 +1 +0 +0   0 0
 P     x   x     2   y-1   1   1     0       1.0 -1 +0 +1   0 0
 P     x   x     2   y-1   z   z     0       1.0 -1 +0 -1   0 0
 Vx    x   x     0   y-1   0   z-1   0       0   0 0
 Vy    x-1 x-1   1   y-1   0   z-1   0       1.0 -1 +0 +0   0 0
 Vz    x-1 x-1   0   y-1   1   z-1   0       1.0 -1 +0 +0   0 0
 T     x   x     1   y-1   0   z     0       1.0 -1 +0 +0   0 0
 Vx    1   x-1   0   y-1   0   0     0       1.0 +0 +0 +1   0 0
 Vy    0   x-1   1   y-1   0   0     0       1.0 +0 +0 +1   0 0
 Vz    0   x-1   0   y-1   0   0     0       0   0 0
 T     1   x-1   1   y-1   0   0     0       1.0 +0 +0 +1   0 0
 Vx    1   x-1   0   y-1   z-1 z-1   0       1.0 +0 +0 -1   0 0
 Vy    0   x-1   1   y-1   z-1 z-1   0       1.0 +0 +0 -1   0 0
 Vz    0   x-1   0   y-1   z   z     0       0   0 0
 T     1   x-1   1   y-1   z   z     0       1.0 +0 +0 -1   0 0
"""


def g16_deck():
    """init.t3c text of the g16 run (own values in the reference's layout)"""
    lines = ["/g16_generator_deck_(own_values)",
             "17-xnumx", "13-ynumy", "15-znumz", "2-mnumx", "2-mnumy", "2-mnumz",
             "2.40000000e+05-xsize(m)", "1.20000000e+05-ysize(m)", "1.80000000e+05-zsize(m)",
             "2-pxinit", "1-pyinit", "2-pzinit", "0.0-pinit(Pa)",
             "0.0-GXKOEF", "9.80665-GYKOEF", "0.0-GZKOEF", "0.0-timesum(Years)",
             "0-nonstab 0.5-xnonstab 0.5-ynonstab 0.5-znonstab",
             "/input_file", "0", "/output_file", "g16.prn b", "/ROCKS"]
    for t in sorted(G16_ROCKS):
        lines.append(f"/rock_{t}")
        lines.append(f"{t} {G16_ROCKS[t]}")
    lines += ["~", "/BONDARIES", G16_BC.rstrip("\n"), "~", "/ROCK_BOXES"]
    lines += G16_RBOX
    lines += ["~", "/T_BOXES"] + G16_TBOX + ["~", ""]
    return "\n".join(lines)


def g16(R=None):
    """The reference's initial-condition generator in3mg (its own program,
    built by oracle/Makefile target in3mg) on the deck g16_init.t3c: rock
    bodies painted on the regular marker lattice (hexahedra with bilinear
    faces, ellipsoid and cylinder shells, immobile markers), temperature
    boxes (trilinear, half-space cooling, transition, sphere, cylinder),
    marker temperatures interpolated from them, two ronurecalc passes, then
    its savers: the .prn state (parsed into g16_generator.npz) and the
    ParaView nodal and composition files (kept as written)."""
    import shutil
    import subprocess
    import tempfile
    deck = g16_deck()
    with open(os.path.join(HERE, "g16_init.t3c"), "w") as f:
        f.write(deck)
    mode = open(os.path.join(HERE, "g13_mode.t3c")).read()
    mode = mode.replace("1.1e4-eroslev 1.2e4-sedilev 1.3e4-waterlev", "-1e30-eroslev 1e30-sedilev -1-waterlev")
    assert "-1e30-eroslev" in mode
    exe = os.path.join(ROOT, "oracle", "_ref", "in3mg")
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "init.t3c"), "w") as f:
            f.write(deck)
        with open(os.path.join(tmp, "mode.t3c"), "w") as f:
            f.write(mode)
        shutil.copy(os.path.join(HERE, "g13_file.t3c"), os.path.join(tmp, "file.t3c"))
        r = subprocess.run([exe], cwd=tmp, stdin=subprocess.DEVNULL, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        files = sorted(os.listdir(tmp))
        print("in3mg wrote", files)
        d = I.read_prn(os.path.join(tmp, "g16.prn"))
        shutil.copy(os.path.join(tmp, "g16.vtr"), os.path.join(HERE, "g16_nodal.vtr"))
        shutil.copy(os.path.join(tmp, "comp_g16.vtr"), os.path.join(HERE, "g16_comp.vtr"))
        with open(os.path.join(HERE, "g16_mode.t3c"), "w") as f:
            f.write(mode)
    keep = {k: v for k, v in d.items() if isinstance(v, np.ndarray) and k not in ("bond_at", "bond_rec", "bond_partner",
                                                                                  "bond_v")}
    scal = {k: v for k, v in d.items() if not isinstance(v, np.ndarray)}
    np.savez_compressed(os.path.join(HERE, "g16_generator.npz"), **keep,
                        **{f"h_{k}": np.asarray(v) for k, v in scal.items()})
    print("g16 written:", d["marknum"], "markers", np.bincount(d["markt"].astype(np.uint8)))


if __name__ == "__main__":
    R = Ref()
    R.lib.refh_set_threads(1)
    which = sys.argv[1:] or ["g1", "g3", "g5", "g6", "g7", "g8", "g9", "g10", "g11", "g12", "g3c2", "g13", "g14",
                             "g15", "g16", "g17"]
    if "g1" in which:
        g1(R)
    if "g3" in which:
        g3(R, 17)
        g3(R, 25)
    if "g5" in which:
        g5(R)
    if "g6" in which:
        g6(R)
    if "g7" in which:
        g7(R)
    if "g8" in which:
        g8(R)
    if "g9" in which:
        g9(R)
    if "g10" in which:
        g10(R)
    if "g11" in which:
        g11(R)
    if "g12" in which:
        g12(R)
    if "g13" in which:
        g13(R)
    if "g3c2" in which:
        g3(R, 25, which=2, seed=2)
        g3(R, 33, which=2, seed=2)
    if "g14" in which:
        g14(R)
    if "g15" in which:
        g15(R)
    if "g16" in which:
        g16(R)
    if "g17" in which:
        g17(R)
