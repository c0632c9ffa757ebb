"""Oracle — TEST INFRASTRUCTURE ONLY.

A plain-C FP64 restatement of the reference Stokes operator, transfers and
norms (``oracle/i3g_oracle.c``), loaded through ctypes.  Only ``tests/``
and ``__graft_entry__.smoke()`` (to check results) and ``bench.py``'s
CPU-baseline leg (the timed CPU side) may use it: the product path never
calls into this package.

Vectors use the reference's ``sol0`` layout (node-major, 4 unknowns per node:
P, vx, vy, vz).  Boundary tables use the reference's global layout
(``bondm`` over every level, ``bondn`` = partner + 1 as a global index).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    """Compile liboracle.so (and, where the reference checkout exists, the
    reference harness under oracle/_ref)."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-s", "liboracle.so"], check=True)
    if os.path.isfile("/root/reference/head3mg.c"):
        subprocess.run(["make", "-C", _HERE, "-s", "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.isfile(path):
            build()
        _LIB = C.CDLL(path)
    return _LIB


class Level(C.Structure):
    _fields_ = [("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
                ("mgp", C.c_long), ("level", C.c_int),
                ("gx", C.c_void_p), ("gy", C.c_void_p), ("gz", C.c_void_p)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def level(nx, ny, nz, mgp, lev, gx, gy, gz):
    """Level descriptor; keeps references to the coordinate arrays."""
    gx = np.ascontiguousarray(gx, np.float64)
    gy = np.ascontiguousarray(gy, np.float64)
    gz = np.ascontiguousarray(gz, np.float64)
    L = Level(nx, ny, nz, mgp, lev, _p(gx), _p(gy), _p(gz))
    L._keep = (gx, gy, gz)
    return L


def eta(L, nu, nxx, nxy, nxz, nyz, viscmod, nubeg, nuend):
    N = L.nx * L.ny * L.nz
    out = [np.zeros(N) for _ in range(4)]
    lib().or_eta(C.byref(L), _p(nu), _p(nxx), _p(nxy), _p(nxz), _p(nyz),
                 C.c_int(viscmod), C.c_double(nubeg), C.c_double(nuend),
                 *[_p(o) for o in out])
    return out


def rhs(L, ro, g, bondm, bondv):
    N = L.nx * L.ny * L.nz
    b = np.zeros(4 * N)
    g = np.asarray(g, np.float64)
    lib().or_rhs(C.byref(L), _p(ro), _p(g), _p(bondm), _p(bondv), _p(b))
    return b


def residual(L, etas, bondm, bondv, bondn, b, x):
    r = np.zeros_like(b)
    lib().or_residual(C.byref(L), *[_p(e) for e in etas], _p(bondm), _p(bondv),
                      _p(bondn), _p(b), _p(x), _p(r))
    return r


def wnorms(L, etaC, bondm, b, r):
    out = np.zeros(2)
    lib().or_wnorms(C.byref(L), _p(etaC), _p(bondm), _p(b), _p(r), _p(out))
    return out


def stress(L, etas, vx, vy, vz):
    N = L.nx * L.ny * L.nz
    eps = np.zeros(6 * N, np.float32)
    sig = np.zeros(6 * N, np.float32)
    lib().or_stress(C.byref(L), *[_p(e) for e in etas], _p(vx), _p(vy), _p(vz),
                    _p(eps), _p(sig))
    return eps.reshape(6, N), sig.reshape(6, N)


def ronurestrict(F, Cl, nu_f, viscmod=0):
    nu_c = np.zeros(Cl.nx * Cl.ny * Cl.nz, np.float32)
    lib().or_ronurestrict(C.byref(F), C.byref(Cl), _p(nu_f), _p(nu_c), C.c_int(viscmod))
    return nu_c


def vprestrict(F, Cl, bondm, r_f, nd_f, nd_c):
    b_c = np.zeros(4 * Cl.nx * Cl.ny * Cl.nz)
    lib().or_vprestrict(C.byref(F), C.byref(Cl), _p(bondm), _p(r_f), _p(nd_f),
                        _p(nd_c), _p(b_c))
    return b_c


def vpprolong(F, Cl, x_c, x_f, v1koef, p1koef):
    x_f = np.array(x_f, np.float64, copy=True)
    lib().or_vpprolong(C.byref(F), C.byref(Cl), _p(x_c), _p(x_f),
                       C.c_double(v1koef), C.c_double(p1koef))
    return x_f


def gs_sweep(L, etas, bondm, bondv, bondn, b, x, p0koef, v0koef):
    """one lexicographic gausmat4 sweep (move3mg.c:3087-3131), interleaved x"""
    x = np.array(x, np.float64, copy=True)
    lib().or_gs_sweep(C.byref(L), *[_p(e) for e in etas], _p(bondm), _p(bondv),
                      _p(bondn), _p(b), _p(x), C.c_double(p0koef), C.c_double(v0koef))
    return x


def colour_phase(L, etas, bondm, b, x, phase, p0koef, v0koef):
    """one phase of the colour-ordered sweep (gausmat4's update rules on a
    colour of one velocity component, or on every P row), interleaved x"""
    x = np.array(x, np.float64, copy=True)
    lib().or_colour_phase(C.byref(L), *[_p(e) for e in etas], _p(bondm), _p(b), _p(x), C.c_int(phase),
                          C.c_double(p0koef), C.c_double(v0koef))
    return x


# ---- markers ----------------------------------------------------------------

ROCK_FIELDS = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht")


class Rock(C.Structure):
    _fields_ = [(f, C.c_double) for f in ROCK_FIELDS]


def rocks_array(rocks):
    """rocks: list of dicts with ROCK_FIELDS"""
    arr = (Rock * len(rocks))()
    for i, r in enumerate(rocks):
        for f in ROCK_FIELDS:
            setattr(arr[i], f, float(r.get(f, 0.0)))
    return arr


def mk_cells(x, y, z, gx, gy, gz):
    n = len(x)
    out = np.zeros(3 * n, np.int32)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    lib().or_mk_cells(C.c_long(n), _p(x), _p(y), _p(z), _p(gx), C.c_long(len(gx)), _p(gy),
                      C.c_long(len(gy)), _p(gz), C.c_long(len(gz)), _p(out))
    return out.reshape(n, 3)


def ronurecalc(x, y, z, tk, eta, mtype, gx, gy, gz, pr, rocks, fld, nubeg, nuend, nukoef=0.0,
               densimod=1, timestep=1.0, vdeep=1e30, nudeep=1e30):
    """Returns (fld (11,N) float32, eta float32) after the reference's
    marker -> node transfer; inputs are not modified."""
    fld = np.array(fld, np.float32, copy=True)
    eta = np.array(eta, np.float32, copy=True)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    ra = rocks_array(rocks)
    lib().or_ronurecalc(C.c_long(len(x)), _p(x), _p(y), _p(z), _p(tk), _p(eta), _p(mtype), _p(gx),
                        C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)), _p(pr), ra,
                        C.c_double(nubeg), C.c_double(nuend), C.c_double(nukoef), C.c_int(densimod),
                        C.c_double(timestep), C.c_double(vdeep), C.c_double(nudeep), _p(fld))
    return fld, eta


def vxyz(vx, vy, vz, gx, gy, gz, x, y, z):
    v = np.zeros(3)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    lib().or_vxyz(_p(vx), _p(vy), _p(vz), _p(gx), C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz),
                  C.c_long(len(gz)), C.c_double(x), C.c_double(y), C.c_double(z), _p(v))
    return v


def epscalc(st12, pr, gx, gy, gz, x, y, z):
    """epscalc1 at a point: [exx eyy ezz exy exz eyz, P, sxx .. syz, dissipation]"""
    st = [np.ascontiguousarray(a, np.float32) for a in st12]
    sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    prd = np.ascontiguousarray(pr, np.float64)
    out = np.zeros(14)
    lib().or_epscalc(sp, _p(prd), _p(gx), C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)),
                     C.c_double(x), C.c_double(y), C.c_double(z), _p(out))
    return out


def mk_props(x, y, z, mtype, tk, eta, e, rocks, st12, pr, gx, gy, gz, nubeg, nuend, dt, healrate):
    """movemark's (eta, e) update of the Newtonian markers; returns copies"""
    eta = np.array(eta, np.float32, copy=True)
    e = np.array(e, np.float32, copy=True)
    st = [np.ascontiguousarray(a, np.float32) for a in st12]
    sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    f = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    t = np.ascontiguousarray(mtype, np.int8)
    tkf = np.ascontiguousarray(tk, np.float32)
    prd = np.ascontiguousarray(pr, np.float64)
    ra = rocks_array(rocks)
    lib().or_mk_props(C.c_long(len(x)), *[_p(a) for a in f], _p(t), _p(tkf), _p(eta), _p(e), ra, sp, _p(prd), _p(gx),
                      C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)), C.c_double(nubeg),
                      C.c_double(nuend), C.c_double(dt), C.c_double(healrate))
    return eta, e


def rk4(x, y, z, tk, eta, mtype, vx, vy, vz, gx, gy, gz, dt, depth=None):
    """RK4 move; depth=(zdeep, vdeep, tdeep) also applies movemark's depth
    rules and returns the new types as a sixth value"""
    x, y, z, tk, eta = (np.array(a, np.float32, copy=True) for a in (x, y, z, tk, eta))
    t = np.array(mtype, np.int8, copy=True)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    zd, vd, td = depth if depth is not None else (1e30, 1e30, 0.0)
    v = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
    lib().or_rk4(C.c_long(len(x)), _p(x), _p(y), _p(z), _p(tk), _p(eta), _p(t), *[_p(a) for a in v],
                 _p(gx), C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)),
                 C.c_double(dt), C.c_double(zd), C.c_double(vd), C.c_double(td))
    if depth is not None:
        return x, y, z, tk, eta, t
    return x, y, z, tk, eta


def tkrecalc(tk, bondm, bondv, bondn):
    tk = np.array(tk, np.float32, copy=True)
    lib().or_tkrecalc(C.c_long(len(tk)), _p(bondm), _p(bondv), _p(bondn), _p(tk))
    return tk


# ---- heat -------------------------------------------------------------------

class HeatSys(C.Structure):
    _fields_ = [("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
                ("gx", C.c_void_p), ("gy", C.c_void_p), ("gz", C.c_void_p),
                ("ro", C.c_void_p), ("cp", C.c_void_p), ("kt", C.c_void_p), ("ht", C.c_void_p),
                ("tk0", C.c_void_p), ("dt", C.c_double),
                ("bondm_t", C.c_void_p), ("bondv", C.c_void_p), ("bondn", C.c_void_p),
                ("timesum", C.c_double), ("timebeg", C.c_double), ("timeend", C.c_double),
                ("tempend", C.c_double), ("src", C.c_void_p)]


def heat_sys(gx, gy, gz, ro, cp, kt, ht, tk0, dt, bondm, bondv, bondn, timesum=0.0, timebeg=1e30,
             timeend=1e30, tempend=0.0, src=None):
    """Level-0 heat system descriptor; bondm may be the full reference table
    (its T block at 8*nodenum is taken) or the T block itself.  src: the
    heaterr source per node (heat_source), default the radiogenic ht."""
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    N = len(gx) * len(gy) * len(gz)
    bondm = np.asarray(bondm)
    bt = np.ascontiguousarray(bondm[8 * N:9 * N] if len(bondm) > N else bondm, np.int32)
    keep = [gx, gy, gz] + [np.ascontiguousarray(a, np.float32) for a in (ro, cp, kt, ht, tk0)] + [
        bt, np.ascontiguousarray(bondv, np.float32), np.ascontiguousarray(bondn, np.int32)]
    H = HeatSys(len(gx), len(gy), len(gz), *[_p(a) for a in keep[:3]], *[_p(a) for a in keep[3:8]], dt,
                *[_p(a) for a in keep[8:]], timesum, timebeg, timeend, tempend)
    if src is not None:
        src = np.ascontiguousarray(src, np.float32)
        keep.append(src)
        H.src = src.ctypes.data
    H._keep = keep
    H.N = N
    return H


def heat_source(gx, gy, gz, ro, cp, ht, tk0, et, vx, vy, vz, pr, s6, e6, bondm, adiabyn, frictyn,
                gk=(0.0, 9.81, 0.0), ysize=None, stp100=0.0):
    """(src, rate): heaterr's right-hand side with adiabatic (adiabyn 1/2)
    and shear (frictyn) heating as stored in tk1 (float32), and the double
    rate src/(ro cp) it returns, per level-0 node."""
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    N = len(gx) * len(gy) * len(gz)
    bondm = np.asarray(bondm)
    bt = np.ascontiguousarray(bondm[8 * N:9 * N] if len(bondm) > N else bondm, np.int32)
    f = [np.ascontiguousarray(a, np.float32) for a in (ro, cp, ht, tk0, et, vx, vy, vz)]
    prd = np.ascontiguousarray(pr, np.float64)
    s6 = [np.ascontiguousarray(a, np.float32) for a in s6]
    e6 = [np.ascontiguousarray(a, np.float32) for a in e6]
    sp = (C.c_void_p * 6)(*[a.ctypes.data for a in s6])
    ep = (C.c_void_p * 6)(*[a.ctypes.data for a in e6])
    gkd = np.ascontiguousarray(gk, np.float64)
    src, rate = np.zeros(N, np.float32), np.zeros(N)
    lib().or_heat_source(C.c_long(len(gx)), C.c_long(len(gy)), C.c_long(len(gz)), _p(gx), _p(gy), _p(gz),
                         *[_p(a) for a in f], _p(prd), sp, ep, _p(bt), C.c_int(adiabyn), C.c_int(frictyn),
                         _p(gkd), C.c_double(gy[-1] if ysize is None else ysize), C.c_double(stp100),
                         _p(src), _p(rate))
    return src, rate


def heat_matrix(H):
    """(scipy CSR matrix, rhs) of the level-0 heat system."""
    import scipy.sparse as sp
    N = H.N
    cnt = np.zeros(N, np.int32)
    cols = np.zeros(7 * N, np.int32)
    vals = np.zeros(7 * N)
    rhs = np.zeros(N)
    lib().or_heat_rows(C.byref(H), _p(cnt), _p(cols), _p(vals), _p(rhs))
    mask = np.arange(7)[None, :] < cnt[:, None]
    rows = np.repeat(np.arange(N), 7).reshape(N, 7)
    A = sp.csr_matrix((vals.reshape(N, 7)[mask], (rows[mask], cols.reshape(N, 7)[mask])), shape=(N, N))
    return A, rhs


def heat_maxrate(H):
    lib().or_heat_maxrate.restype = C.c_double
    return lib().or_heat_maxrate(C.byref(H))


def heat_norms(H, x):
    out = np.zeros(2)
    x = np.ascontiguousarray(x, np.float64)
    lib().or_heat_norms(C.byref(H), _p(x), _p(out))
    return out


def mk_tk_update(x, y, z, mtype, markk, gx, gy, gz, tk, tk2, kt, ro, cp, heatdif, dt, zdeep=1e30, tdeep=0.0):
    markk = np.array(markk, np.float32, copy=True)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    arrs = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    nodes = [np.ascontiguousarray(a, np.float32) for a in (tk, tk2, kt, ro, cp)]
    t = np.ascontiguousarray(mtype, np.int8)
    lib().or_mk_tk_update(C.c_long(len(x)), *[_p(a) for a in arrs], _p(t), _p(markk), _p(gx), C.c_long(len(gx)),
                          _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)), *[_p(a) for a in nodes],
                          C.c_double(heatdif), C.c_double(dt), C.c_double(zdeep), C.c_double(tdeep))
    return markk


def err_rows(L, nu, ro, g, nubeg, nuend, st12, pr, vx, vy, vz, bondm, bondv, bondn, sol):
    N = L.nx * L.ny * L.nz
    rows = np.zeros((5, 4 * N, 2))
    g = np.asarray(g, np.float64)
    st12 = np.ascontiguousarray(st12, np.float32)
    lib().or_err_rows(C.byref(L), _p(nu), _p(ro), _p(g), C.c_double(nubeg), C.c_double(nuend), _p(st12),
                      _p(np.ascontiguousarray(pr, np.float64)), _p(vx), _p(vy), _p(vz), _p(bondm), _p(bondv),
                      _p(bondn), _p(np.ascontiguousarray(sol, np.float64)), _p(rows))
    return rows


def native_norms(rows, maxvxyz, bondm=None):
    """vpiterate's stoksum, stoksum1, stoksum2, contsum, contsumb, bondsum
    (move3mg.c:2405-2596) from residual-mode rows; maxvxyz as the reference
    (RMS velocity range, 1e-8 when zero).  bondm splits the continuity rows
    into free (contsum) and BC (contsumb) pressure cells."""
    mom = np.concatenate([rows[c][np.isfinite(rows[c][:, 0])] for c in (1, 2, 3)])
    ival, ival1 = mom[:, 0], mom[:, 1] / maxvxyz
    ok = np.isfinite(rows[0][0::4, 0])     # continuity rows live at u = 4 n
    cont = rows[0][0::4, 0] / maxvxyz
    out = dict(stoksum=np.sqrt(np.mean(ival ** 2)), stoksum1=np.sqrt(np.mean(ival1 ** 2)),
               stoksum2=np.sqrt(np.mean(ival1 ** 2 * ival ** 2)),
               bondsum=np.sqrt(np.mean(rows[4][np.isfinite(rows[4][:, 0]), 0] ** 2)), cont=cont[ok])
    if bondm is not None:
        bc = np.asarray(bondm)[:len(cont)] != 0
        out["contsum"] = np.sqrt(np.mean(cont[ok & ~bc] ** 2))
        out["contsumb"] = np.sqrt(np.mean(cont[ok & bc] ** 2)) if np.any(ok & bc) else 0.0
    return out


def maxvxyz(vx, vy, vz, nx, ny, nz, divvmax=0.0):
    """RMS of the float velocity ranges over existing unknowns
    (move3mg.c:1648-1672, 1766-1770)"""
    m1 = (np.arange(nx * ny * nz) // ny) % nx
    m2 = np.arange(nx * ny * nz) % ny
    m3 = np.arange(nx * ny * nz) // (nx * ny)
    sel = [(m2 < ny - 1) & (m3 < nz - 1), (m1 < nx - 1) & (m3 < nz - 1), (m1 < nx - 1) & (m2 < ny - 1)]
    d = [float(np.float32(v)[s].max()) - float(np.float32(v)[s].min()) for v, s in zip((vx, vy, vz), sel)]
    mv = np.sqrt(sum(x * x for x in d) / 3.0)
    if mv <= 0:
        mv = 1e-8
    if divvmax > 0 and mv > divvmax:
        mv = divvmax
    return mv


# ---- full rheology -----------------------------------------------------------

ROCKX_FIELDS = ROCK_FIELDS + ("ss", "mm", "ll", "s1", "a0", "a1", "b0", "b1", "e0", "e1")


class RockX(C.Structure):
    _fields_ = [(f, C.c_double) for f in ROCKX_FIELDS]


class VisPrm(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("nubeg", "nuend", "timestep", "timesum", "convend")]


class MkX(C.Structure):
    _fields_ = [("nukoef", C.c_double), ("densimod", C.c_int), ("vdeep", C.c_double), ("nudeep", C.c_double),
                ("eroslev", C.c_double), ("sedilev", C.c_double), ("waterlev", C.c_double), ("sedtype", C.c_int),
                ("mnumy", C.c_long), ("teclmin", C.c_double), ("teclmax", C.c_double), ("maxpmelt", C.c_double)]


class MoveX(C.Structure):
    _fields_ = [("outgrid", C.c_int), ("markmod", C.c_int), ("vyfluid", C.c_double), ("vymelt", C.c_double),
                ("g", C.c_double * 3)]


def rocksx_array(rocks):
    """rocks: list of dicts with ROCKX_FIELDS (missing ones 0)"""
    arr = (RockX * len(rocks))()
    for i, r in enumerate(rocks):
        for f in ROCKX_FIELDS:
            setattr(arr[i], f, float(r.get(f, 0.0)))
    return arr


def _visprm(nubeg, nuend, timestep, timesum, convend):
    return VisPrm(float(nubeg), float(nuend), float(timestep), float(timesum), float(convend))


def viscalc(rock, mm2, mtk, mpb, eii, sii, rii, dis, e, rr, rn, nubeg=1e18, nuend=1e25, timestep=0.0,
            timesum=0.0, convend=0.0):
    """viscalc for one marker: returns (eta, e, rr, rn)"""
    ra = rocksx_array([rock])
    fe, frr, frn = C.c_float(e), C.c_float(rr), C.c_float(rn)
    lib().or_viscalc.restype = C.c_double
    P = _visprm(nubeg, nuend, timestep, timesum, convend)
    v = lib().or_viscalc(ra, C.c_int(mm2), C.c_double(mtk), C.c_double(mpb), C.c_double(eii), C.c_double(sii),
                         C.c_double(rii), C.c_double(dis), C.byref(fe), C.byref(frr), C.byref(frn), C.byref(P))
    return v, fe.value, frr.value, frn.value


def mk_props_x(x, y, z, mtype, tk, eta, e, rr, rn, rocks, st12, pr, gx, gy, gz, nubeg, nuend, dt, timesum,
               convend, healrate, outgrid=0):
    """movemark's property update with the full viscalc; returns copies of
    (eta, e, rr, rn)"""
    eta, e, rr, rn = (np.array(a, np.float32, copy=True) for a in (eta, e, rr, rn))
    st = [np.ascontiguousarray(a, np.float32) for a in st12]
    sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    f = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    t = np.ascontiguousarray(mtype, np.int8)
    tkf = np.ascontiguousarray(tk, np.float32)
    prd = np.ascontiguousarray(pr, np.float64)
    ra = rocksx_array(rocks)
    P = _visprm(nubeg, nuend, dt, timesum, convend)
    lib().or_mk_props_x(C.c_long(len(x)), *[_p(a) for a in f], _p(t), _p(tkf), _p(eta), _p(e), _p(rr), _p(rn), ra,
                        sp, _p(prd), _p(gx), C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz),
                        C.c_long(len(gz)), C.byref(P), C.c_double(healrate), C.c_int(outgrid))
    return eta, e, rr, rn


def ronurecalc_x(x, y, z, tk, eta, mtype, e, rr, rn, gx, gy, gz, pr, st12, rocks, fld, nubeg, nuend, timestep,
                 timesum, convend, nukoef=0.0, densimod=1, vdeep=1e30, nudeep=1e30, eroslev=-1e30, sedilev=1e30,
                 waterlev=-1.0, sedtype=3, mnumy=2, teclmin=0.0, teclmax=0.0, maxpmelt=1e10, markw=None,
                 markex=None):
    """ronurecalc with the full rheology, melt types (21-49) from their
    water content markw and extracted melt markex (None: zero); returns
    (fld, eta, type, e, rr, rn) copies; raises on densimod >= 2"""
    fld = np.array(fld, np.float32, copy=True)
    eta, e, rr, rn = (np.array(a, np.float32, copy=True) for a in (eta, e, rr, rn))
    t = np.array(mtype, np.int8, copy=True)
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    f = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, tk)]
    if st12 is not None:
        st = [np.ascontiguousarray(a, np.float32) for a in st12]
        sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
    else:
        sp = None
    prd = np.ascontiguousarray(pr, np.float64)
    ra = rocksx_array(rocks)
    P = _visprm(nubeg, nuend, timestep, timesum, convend)
    M = MkX(float(nukoef), int(densimod), float(vdeep), float(nudeep), float(eroslev), float(sedilev),
            float(waterlev), int(sedtype), int(mnumy), float(teclmin), float(teclmax), float(maxpmelt))
    mw = None if markw is None else np.ascontiguousarray(markw, np.float32)
    mx = None if markex is None else np.ascontiguousarray(markex, np.float32)
    rc = lib().or_ronurecalc_x(C.c_long(len(x)), *[_p(a) for a in f], _p(eta), _p(t), _p(e), _p(rr), _p(rn),
                               _p(gx), C.c_long(len(gx)), _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)),
                               _p(prd), sp, ra, C.byref(P), C.byref(M), _p(mw) if mw is not None else None,
                               _p(mx) if mx is not None else None, _p(fld))
    if rc:
        raise ValueError("a participating marker needs densimod >= 2")
    return fld, eta, t, e, rr, rn


def rk4_x(x, y, z, tk, eta, mtype, md, vx, vy, vz, pr, gx, gy, gz, dt, depth=(1e30, 1e30, 0.0), outgrid=0,
          markmod=0, vyfluid=0.0, vymelt=0.0, g=(0.0, 9.81, 0.0)):
    """movemark's motion with outgrid, markmod, fluid and immobile markers;
    returns (x, y, z, tk, eta, type) copies"""
    x, y, z, tk, eta = (np.array(a, np.float32, copy=True) for a in (x, y, z, tk, eta))
    t = np.array(mtype, np.int8, copy=True)
    mdp = np.ascontiguousarray(md, np.float32) if md is not None else None
    gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
    v = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
    prd = np.ascontiguousarray(pr, np.float64)
    M = MoveX(int(outgrid), int(markmod), float(vyfluid), float(vymelt), (C.c_double * 3)(*g))
    lib().or_rk4_x(C.c_long(len(x)), _p(x), _p(y), _p(z), _p(tk), _p(eta), _p(t),
                   _p(mdp) if mdp is not None else None, *[_p(a) for a in v], _p(prd), _p(gx), C.c_long(len(gx)),
                   _p(gy), C.c_long(len(gy)), _p(gz), C.c_long(len(gz)), C.c_double(dt), C.c_double(depth[0]),
                   C.c_double(depth[1]), C.c_double(depth[2]), C.byref(M))
    return x, y, z, tk, eta, t


def meltpart11(mm2, mtk, mpb, mwa=0.0, maxpmelt=1e10):
    """melt fraction and latent heat (J/kg) of rock type mm2 at mtk (K),
    mpb (bar) with water content mwa"""
    xm, hl = C.c_double(), C.c_double()
    lib().or_meltpart11(C.c_int(mm2), C.c_double(mtk), C.c_double(mpb), C.c_double(mwa), C.c_double(maxpmelt),
                        C.byref(xm), C.byref(hl))
    return xm.value, hl.value


def meltpart0(rocks, mm2, mtk, mpb, mwa=0.0, mex=0.0, densimod=1, teclmin=0.0, teclmax=0.0, maxpmelt=1e10):
    """eps1[20..26] of a marker of melt type mm2 and the return flag"""
    e = (C.c_double * 7)()
    rc = lib().or_meltpart0(rocksx_array(rocks), C.c_int(mm2), C.c_double(mtk), C.c_double(mpb), C.c_double(mwa),
                            C.c_double(mex), C.c_int(densimod), C.c_double(teclmin), C.c_double(teclmax),
                            C.c_double(maxpmelt), e)
    return list(e), rc


def set_sum64(on):
    """marker -> node sums of ronurecalc / ronurecalc_x in double (True) or
    in float marker by marker as the reference (False, the default)"""
    lib().or_set_sum64(C.c_int(1 if on else 0))
