"""i3elvis_amd — MI355X-native hot path of I3ELVIS-style staggered-grid
marker-in-cell geodynamics.

The product is the C ABI in ``include/i3g.h`` implemented by the in-tree
library ``i3elvis_amd/libi3g.so`` (hand-written gfx950 HIP kernels + C host
layer).  This module is a thin ctypes view of that ABI used by tests and the
benchmark; it never computes anything itself and raises if the library is
missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libi3g.so")
MAXLEV = 10
_LIB = None


class I3GError(RuntimeError):
    pass


def lib():
    """The native library; raises I3GError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.isfile(LIBPATH):
            raise I3GError("libi3g.so not built: run python i3elvis_amd/build.py")
        _LIB = C.CDLL(LIBPATH)
        _LIB.i3g_bc_expand.restype = C.c_long
        _LIB.i3g_error_string.restype = C.c_char_p
    return _LIB


def check(rc, what=""):
    if rc < 0:
        msg = lib().i3g_error_string(rc).decode()
        raise I3GError(f"{what}: {msg} ({rc})")
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Hier(C.Structure):
    _fields_ = ([("nlev", C.c_int)] +
                [(n, C.c_long * MAXLEV) for n in ("mgx", "mgy", "mgz", "mgp", "mgn", "mgs")] +
                [("ntot", C.c_long)])


class BcRule(C.Structure):
    _fields_ = ([("var", C.c_int)] +
                [(n, C.c_long) for n in ("m10", "m11", "m20", "m21", "m30", "m31")] +
                [(n, C.c_int) for n in ("rel10", "rel11", "rel20", "rel21", "rel30", "rel31")] +
                [("cnst", C.c_double), ("koef", C.c_double),
                 ("dm1", C.c_int), ("dm2", C.c_int), ("dm3", C.c_int)])


class Setup(C.Structure):
    _fields_ = [("nx", C.c_long), ("ny", C.c_long), ("nz", C.c_long),
                ("gx", C.c_void_p), ("gy", C.c_void_p), ("gz", C.c_void_p),
                ("multinum", C.c_int), ("bondm", C.c_void_p), ("bondv", C.c_void_p),
                ("bondn", C.c_void_p), ("nrec", C.c_long)]


class StokesParams(C.Structure):
    _fields_ = [("rtol", C.c_double), ("precision", C.c_int), ("max_outer", C.c_int),
                ("krylov", C.c_int), ("max_vcycles", C.c_int), ("nu_pre", C.c_int * MAXLEV), ("nu_post", C.c_int * MAXLEV), ("nu_coarse", C.c_int), ("inner_m", C.c_int), ("inner_tol", C.c_double),
                ("p0koef", C.c_double), ("p1koef", C.c_double), ("v0koef", C.c_double),
                ("v1koef", C.c_double), ("inner_rtol", C.c_double), ("warm_start", C.c_int),
                ("pxinit", C.c_long), ("pyinit", C.c_long), ("pzinit", C.c_long),
                ("pinit", C.c_double), ("verbose", C.c_int)]


class StokesStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("outer_iters", C.c_int), ("vcycles", C.c_int),
                ("rel_res", C.c_double), ("res0", C.c_double), ("seconds", C.c_double),
                ("hist", C.c_double * 256), ("nhist", C.c_int)]


def hierarchy(nx, ny, nz, multinum):
    h = Hier()
    check(lib().i3g_hierarchy(C.c_long(nx), C.c_long(ny), C.c_long(nz), C.c_int(multinum),
                              C.byref(h)), "hierarchy")
    return h


def level_lines(h, l, gx, gy, gz):
    out = [np.zeros(h.mgx[l]), np.zeros(h.mgy[l]), np.zeros(h.mgz[l])]
    check(lib().i3g_level_lines(C.byref(h), C.c_int(l), _p(gx), _p(gy), _p(gz), *[_p(o) for o in out]))
    return out


def freeslip_rules(t_top=273.0, t_bottom=1600.0):
    rules = (BcRule * 64)()
    n = check(lib().i3g_bc_rules_freeslip(C.c_double(t_top), C.c_double(t_bottom), rules, 64))
    return rules, n


def expand_bc(rules, nrules, h):
    """Reference-layout boundary tables (bondm, bondv, bondn) over all levels."""
    nodenum = h.mgx[0] * h.mgy[0] * h.mgz[0]
    cap = 16 * (h.mgx[0] * h.mgy[0] + h.mgy[0] * h.mgz[0] + h.mgx[0] * h.mgz[0]) * h.nlev + 1024
    bondm = np.zeros(10 * nodenum, np.int32)
    bondv = np.zeros(2 * cap, np.float32)
    bondn = np.zeros(cap, np.int32)
    nb = check(lib().i3g_bc_expand(rules, C.c_int(nrules), C.byref(h), _p(bondm), _p(bondv),
                                   _p(bondn), C.c_long(cap)), "bc_expand")
    return bondm, bondv[:2 * (nb + 1)].copy(), bondn[:nb + 1].copy()


def make_config(which, nx, ny, nz, seed=1, multinum=None, t_top=273.0, t_bottom=1600.0):
    """Seeded synthetic configuration (see i3g_gen_config) with free-slip
    boundary tables and the multigrid hierarchy."""
    N = nx * ny * nz
    gx, gy, gz = np.zeros(nx), np.zeros(ny), np.zeros(nz)
    f = {k: np.zeros(N, np.float32) for k in ("nu", "ro", "tk", "nxx", "nxy", "nxz", "nyz")}
    par = np.zeros(2)
    check(lib().i3g_gen_config(C.c_int(which), C.c_long(nx), C.c_long(ny), C.c_long(nz),
                               C.c_ulonglong(seed), _p(gx), _p(gy), _p(gz), _p(f["nu"]),
                               _p(f["ro"]), _p(f["tk"]), _p(f["nxx"]), _p(f["nxy"]),
                               _p(f["nxz"]), _p(f["nyz"]), _p(par)), "gen_config")
    if multinum is None:
        multinum = 0
        m = min(nx, ny, nz)
        while multinum < 5 and (m - 5) // 2 + 5 >= 12:
            m = (m - 5) // 2 + 5
            multinum += 1
    h = hierarchy(nx, ny, nz, multinum)
    rules, nr = freeslip_rules(t_top, t_bottom)
    bondm, bondv, bondn = expand_bc(rules, nr, h)
    cfg = dict(which=which, nx=nx, ny=ny, nz=nz, gx=gx, gy=gy, gz=gz, multinum=multinum,
               hier=h, bondm=bondm, bondv=bondv, bondn=bondn, nubeg=par[0], nuend=par[1],
               g=(0.0, 9.81, 0.0), viscmod=0)
    cfg.update(f)
    return cfg


def default_params(**kw):
    p = StokesParams()
    lib().i3g_stokes_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("nu_pre", "nu_post"):
            for i, x in enumerate(v):
                getattr(p, k)[i] = x
        else:
            setattr(p, k, v)
    return p


class Stokes:
    """Device context for one grid (i3g_create .. i3g_destroy)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.N = cfg["nx"] * cfg["ny"] * cfg["nz"]
        self._keep = [np.ascontiguousarray(cfg[k]) for k in ("gx", "gy", "gz", "bondm", "bondv", "bondn")]
        gx, gy, gz, bm, bv, bn = self._keep
        s = Setup(cfg["nx"], cfg["ny"], cfg["nz"], _p(gx), _p(gy), _p(gz), cfg["multinum"],
                  _p(bm), _p(bv), _p(bn), len(bn))
        self.ctx = C.c_void_p()
        check(lib().i3g_create(C.byref(s), C.byref(self.ctx)), "create")

    def close(self):
        if self.ctx:
            lib().i3g_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_coefficients(self, nu=None, nxx=None, nxy=None, nxz=None, nyz=None, ro=None):
        c = self.cfg
        arrs = [np.ascontiguousarray(a if a is not None else c[k], np.float32)
                for a, k in ((nu, "nu"), (nxx, "nxx"), (nxy, "nxy"), (nxz, "nxz"), (nyz, "nyz"), (ro, "ro"))]
        phys = np.array([c["g"][0], c["g"][1], c["g"][2], c["nubeg"], c["nuend"], c["viscmod"]], np.float64)
        check(lib().i3g_set_coefficients(self.ctx, *[_p(a) for a in arrs], _p(phys)), "set_coefficients")

    def residual(self, P, vx, vy, vz):
        P, vx, vy, vz = (np.ascontiguousarray(a, np.float64) for a in (P, vx, vy, vz))
        r = [np.zeros(self.N) for _ in range(4)]
        norms = np.zeros(2)
        check(lib().i3g_stokes_residual(self.ctx, _p(P), _p(vx), _p(vy), _p(vz), *[_p(a) for a in r],
                                        _p(norms)), "residual")
        return r, norms

    def set_solution(self, P, vx, vy, vz):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (P, vx, vy, vz)]
        check(lib().i3g_stokes_set_solution(self.ctx, *[_p(a) for a in arrs]), "set_solution")

    def get_solution(self):
        out = [np.zeros(self.N) for _ in range(4)]
        check(lib().i3g_stokes_get_solution(self.ctx, *[_p(a) for a in out]), "get_solution")
        return out

    def solve(self, params=None, check_conv=True, **kw):
        p = params or default_params(**kw)
        st = StokesStats()
        rc = lib().i3g_stokes_solve_resident(self.ctx, C.byref(p), C.byref(st))
        if rc != -4 or check_conv:
            check(rc, "stokes_solve")
        return st

    def combine_start(self, base, dirs):
        """resident iterate = base + sum a_i dirs[i] with residual-optimal a;
        vectors are [P, vx, vy, vz] lists whose entries may be None.
        Returns (a, rel)."""
        vecs = [base] + list(dirs)
        keep = [None if p is None else np.ascontiguousarray(p, np.float64) for v in vecs for p in v]
        arr = (C.c_void_p * len(keep))(*[None if p is None else p.ctypes.data for p in keep])
        a = np.zeros(max(1, len(dirs)))
        rel = C.c_double()
        check(lib().i3g_stokes_combine_start(self.ctx, C.c_int(len(dirs)), arr, _p(a), C.byref(rel)),
              "combine_start")
        return a[:len(dirs)], rel.value

    def learned_start(self, net, pscale, vscale, gauge=(2, 1, 2)):
        """i3g_stokes_learned_start: the learned start built on the device
        (features, inference, lithostatic pressure, residual-optimal
        combination, zero-start fallback).  Returns the info dict."""
        info = np.zeros(6)
        g = (C.c_long * 3)(*[int(v) for v in gauge])
        check(lib().i3g_stokes_learned_start(self.ctx, net.h, C.c_double(pscale), C.c_double(vscale), g, _p(info)),
              "stokes_learned_start")
        return dict(coef_dP=float(info[0]), coef_v=float(info[1]), rel_learned=float(info[2]),
                    rel_lithostatic=float(info[3]), rel_fallback=float(info[4]),
                    start="network" if info[5] else "fallback")

    def level_nodes(self, level):
        c = self.cfg
        h = c.get("hier") or hierarchy(c["nx"], c["ny"], c["nz"], c["multinum"])
        return int(h.mgx[level]) * int(h.mgy[level]) * int(h.mgz[level])

    def restrict(self, level, fine):
        """vprestrict(level): fine residual planes [P, vx, vy, vz] of
        level-1 -> coarse right-hand side planes of `level` (FP64)"""
        keep = [np.ascontiguousarray(a, np.float64) for a in fine]
        out = [np.zeros(self.level_nodes(level)) for _ in range(4)]
        fin = (C.c_void_p * 4)(*[a.ctypes.data for a in keep])
        cou = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
        check(lib().i3g_stokes_restrict(self.ctx, C.c_int(level), fin, cou), "stokes_restrict")
        return out

    def prolong(self, level, coarse, fine, v1koef, p1koef):
        """vpprolong(level, 0): returns fine + the weighted coarse correction
        (planes [P, vx, vy, vz], FP64)"""
        keep = [np.ascontiguousarray(a, np.float64) for a in coarse]
        out = [np.array(a, np.float64, copy=True) for a in fine]
        cou = (C.c_void_p * 4)(*[a.ctypes.data for a in keep])
        fin = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
        check(lib().i3g_stokes_prolong(self.ctx, C.c_int(level), cou, fin, C.c_double(v1koef),
                                       C.c_double(p1koef)), "stokes_prolong")
        return out

    def fields(self):
        """(vx, vy, vz float32, pr float64, vmax[3]) of the resident solution
        (i3g_stokes_fields)"""
        out = [np.zeros(self.N, np.float32) for _ in range(3)] + [np.zeros(self.N)]
        vmax = np.zeros(3)
        check(lib().i3g_stokes_fields(self.ctx, *[_p(a) for a in out], _p(vmax)), "stokes_fields")
        return out[0], out[1], out[2], out[3], vmax

    def smooth_phase(self, level, phase, precision, x, b, v0koef, p0koef):
        """one colour phase of the smoother on planes [P, vx, vy, vz] of
        `level` (i3g_stokes_smooth_phase); returns the new planes"""
        out = [np.array(a, np.float64, copy=True) for a in x]
        keep = [np.ascontiguousarray(a, np.float64) for a in b]
        xp = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
        bp = (C.c_void_p * 4)(*[a.ctypes.data for a in keep])
        check(lib().i3g_stokes_smooth_phase(self.ctx, C.c_int(level), C.c_int(phase), C.c_int(precision), xp, bp,
                                            C.c_double(v0koef), C.c_double(p0koef)), "stokes_smooth_phase")
        return out

    def smooth(self, level, precision, nsweeps, x, b, v0koef, p0koef, variant=0):
        """nsweeps smoother sweeps (i3g_stokes_smooth; variant 0 the
        context's default, 1 the seven phase launches, 2 one launch per
        velocity component, 3 the seven phases in z-slabs, 4 the whole FP32
        sweep in one launch on any level, 5 the FP32 sweep in three launches,
        one per velocity component, on any level)"""
        out = [np.array(a, np.float64, copy=True) for a in x]
        keep = [np.ascontiguousarray(a, np.float64) for a in b]
        xp = (C.c_void_p * 4)(*[a.ctypes.data for a in out])
        bp = (C.c_void_p * 4)(*[a.ctypes.data for a in keep])
        check(lib().i3g_stokes_smooth(self.ctx, C.c_int(level), C.c_int(precision), C.c_int(nsweeps),
                                      C.c_int(variant), xp, bp, C.c_double(v0koef), C.c_double(p0koef)),
              "stokes_smooth")
        return out

    def strain_stress(self):
        """(eps, sig): 6 float planes each (xx yy zz xy xz yz) of the resident solution"""
        out = [np.zeros(self.N, np.float32) for _ in range(12)]
        check(lib().i3g_stokes_strain_stress(self.ctx, *[_p(a) for a in out]), "strain_stress")
        return np.stack(out[:6]), np.stack(out[6:])

    def native_norms(self, divvmax=0.0):
        """the reference's stoksum, stoksum1, stoksum2, contsum, contsumb, bondsum, maxvxyz"""
        out = np.zeros(7)
        check(lib().i3g_stokes_native_norms(self.ctx, C.c_double(divvmax), _p(out)), "native_norms")
        return dict(zip(("stoksum", "stoksum1", "stoksum2", "contsum", "contsumb", "bondsum", "maxvxyz"), out))

    def set_timing(self, on=True):
        check(lib().i3g_set_timing(self.ctx, C.c_int(1 if on else 0)))

    def kernel_stats(self):
        cap, nl = 64, 64
        names = C.create_string_buffer(cap * nl)
        ms = np.zeros(cap)
        calls = np.zeros(cap, np.int64)
        by = np.zeros(cap)
        n = check(lib().i3g_kernel_stats(self.ctx, names, C.c_int(nl), _p(ms), _p(calls), _p(by), C.c_int(cap)))
        out = {}
        for q in range(n):
            nm = names.raw[q * nl:(q + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = dict(avg_ms=float(ms[q]), calls=int(calls[q]), bytes=float(by[q]))
        return out


# ---- markers ----------------------------------------------------------------

ROCK_FIELDS = ("nu0", "dh", "dv", "n0", "n1", "ro", "bro", "aro", "cp", "kt", "kf", "kp", "ht",
               "ss", "mm", "ll", "s1", "a0", "a1", "b0", "b1", "e0", "e1")


class Rock(C.Structure):
    _fields_ = [(f, C.c_double) for f in ROCK_FIELDS]


class MkParams(C.Structure):
    """i3g_mk_params; rheology 0 leaves the fields after it unused"""
    _fields_ = [("nubeg", C.c_double), ("nuend", C.c_double), ("nukoef", C.c_double), ("densimod", C.c_int),
                ("timestep", C.c_double), ("vdeep", C.c_double), ("nudeep", C.c_double), ("viscmod", C.c_int),
                ("rheology", C.c_int), ("eroslev", C.c_double), ("sedilev", C.c_double), ("waterlev", C.c_double),
                ("sedtype", C.c_int), ("mnumy", C.c_long), ("teclmin", C.c_double), ("teclmax", C.c_double),
                ("timesum", C.c_double), ("convend", C.c_double), ("outgrid", C.c_int), ("maxpmelt", C.c_double)]


class MoveParams(C.Structure):
    """i3g_move_params"""
    _fields_ = [("outgrid", C.c_int), ("markmod", C.c_int), ("vyfluid", C.c_double), ("vymelt", C.c_double),
                ("g", C.c_double * 3)]


def mk_params(nubeg=1e18, nuend=1e25, nukoef=0.0, densimod=1, timestep=1.0, vdeep=1e30, nudeep=1e30, viscmod=0,
              rheology=0, eroslev=-1e30, sedilev=1e30, waterlev=-1.0, sedtype=3, mnumy=1, teclmin=0.0,
              teclmax=0.0, timesum=0.0, convend=0.0, outgrid=0, maxpmelt=1e10):
    return MkParams(float(nubeg), float(nuend), float(nukoef), int(densimod), float(timestep), float(vdeep),
                    float(nudeep), int(viscmod), int(rheology), float(eroslev), float(sedilev), float(waterlev),
                    int(sedtype), int(mnumy), float(teclmin), float(teclmax), float(timesum), float(convend),
                    int(outgrid), float(maxpmelt))


def rocks_array(rocks):
    arr = (Rock * len(rocks))()
    for i, r in enumerate(rocks):
        vals = r if not isinstance(r, dict) else [r.get(f, 0.0) for f in ROCK_FIELDS]
        for f, v in zip(ROCK_FIELDS, vals):
            setattr(arr[i], f, float(v))
    return arr


class Markers:
    """Device marker set (i3g_markers_create .. destroy)."""

    def __init__(self, gx, gy, gz, capacity, order=None):
        """order: node-average order (i3g_markers_set_order), None keeps the
        library's default (cell-segmented)"""
        self._g = [np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz)]
        self.nx, self.ny, self.nz = (len(a) for a in self._g)
        self.N = self.nx * self.ny * self.nz
        self.h = C.c_void_p()
        check(lib().i3g_markers_create(C.c_long(self.nx), C.c_long(self.ny), C.c_long(self.nz),
                                       *[_p(a) for a in self._g], C.c_long(capacity), C.byref(self.h)),
              "markers_create")
        lib().i3g_markers_count.restype = C.c_long
        if order is not None:
            self.set_order(order)

    def close(self):
        if self.h:
            lib().i3g_markers_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return lib().i3g_markers_count(self.h)

    def last_ms(self):
        """device time of the last operation's kernels"""
        lib().i3g_markers_last_ms.restype = C.c_double
        return lib().i3g_markers_last_ms(self.h)

    def upload(self, x, y, z, tk, eta, mtype, ids=None):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, tk, eta)]
        t = np.ascontiguousarray(mtype, np.int8)
        i = None if ids is None else np.ascontiguousarray(ids, np.int32)
        check(lib().i3g_markers_upload(self.h, C.c_long(len(arrs[0])), *[_p(a) for a in arrs], _p(t), _p(i)),
              "markers_upload")

    def download(self):
        n = self.n
        out = [np.zeros(n, np.float32) for _ in range(5)]
        t = np.zeros(n, np.int8)
        ids = np.zeros(n, np.int32)
        check(lib().i3g_markers_download(self.h, *[_p(a) for a in out], _p(t), _p(ids)), "markers_download")
        return dict(x=out[0], y=out[1], z=out[2], tk=out[3], eta=out[4], type=t, id=ids)

    def cell_index(self):
        n = self.n
        out = [np.zeros(n, np.int32) for _ in range(4)]
        check(lib().i3g_markers_cell_index(self.h, *[_p(a) for a in out]), "markers_cell_index")
        return out

    def sort(self):
        order = np.zeros(self.n, np.int64)
        check(lib().i3g_markers_sort(self.h, _p(order)), "markers_sort")
        return order

    def to_nodes(self, rocks, pr, fld, nubeg, nuend, nukoef=0.0, densimod=1, timestep=1.0, vdeep=1e30,
                 nudeep=1e30, viscmod=0, **rheo):
        """marker -> node transfer; rheo: the i3g_mk_params fields of the
        whole rheology (rheology=1, eroslev, sedilev, waterlev, sedtype,
        mnumy, teclmin, teclmax, timesum, convend, maxpmelt)"""
        fld = np.array(fld, np.float32, copy=True)
        pr = np.ascontiguousarray(pr, np.float64)
        prm = mk_params(nubeg, nuend, nukoef, densimod, timestep, vdeep, nudeep, viscmod, **rheo)
        ra = rocks_array(rocks)
        check(lib().i3g_markers_to_nodes(self.h, ra, C.c_int(len(rocks)), C.byref(prm), _p(pr),
                                         *[_p(fld[q]) for q in range(11)]), "markers_to_nodes")
        return fld

    def set_depth_rules(self, zdeep=1e30, vdeep=1e30, tdeep=0.0):
        check(lib().i3g_markers_set_depth_rules(self.h, C.c_double(zdeep), C.c_double(vdeep), C.c_double(tdeep)),
              "markers_set_depth_rules")

    def set_strain(self, e):
        e = np.ascontiguousarray(e, np.float32)
        check(lib().i3g_markers_set_strain(self.h, _p(e)), "markers_set_strain")

    def get_strain(self):
        e = np.zeros(self.n, np.float32)
        check(lib().i3g_markers_get_strain(self.h, _p(e)), "markers_get_strain")
        return e

    def epscalc(self, st12, pr):
        """(n, 14): eps xx yy zz xy xz yz, P, sigma xx .. yz, dissipation at
        every marker (current order)"""
        st = [np.ascontiguousarray(a, np.float32) for a in st12]
        sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
        prd = np.ascontiguousarray(pr, np.float64)
        out = np.zeros((self.n, 14))
        check(lib().i3g_markers_epscalc(self.h, sp, _p(prd), _p(out)), "markers_epscalc")
        return out

    def update_props(self, rocks, nubeg, nuend, timestep, healrate, st12, pr, rheology=0, timesum=0.0, convend=0.0,
                     outgrid=0):
        st = [np.ascontiguousarray(a, np.float32) for a in st12]
        sp = (C.c_void_p * 12)(*[a.ctypes.data for a in st])
        prd = np.ascontiguousarray(pr, np.float64)
        prm = mk_params(nubeg, nuend, 0.0, 1, timestep, rheology=rheology, timesum=timesum, convend=convend,
                        outgrid=outgrid)
        ra = rocks_array(rocks)
        check(lib().i3g_markers_update_props(self.h, ra, C.c_int(len(rocks)), C.byref(prm), C.c_double(healrate), sp,
                                             _p(prd)), "markers_update_props")

    def set_tbc(self, bondm, bondv, bondn):
        """Temperature boundary rows from the reference-layout tables (bondm
        over all levels or just level 0's nine blocks); None switches off."""
        if bondm is None:
            check(lib().i3g_markers_set_tbc(self.h, None, None, None, C.c_long(0)), "markers_set_tbc")
            return
        bt = np.ascontiguousarray(np.asarray(bondm)[8 * self.N:9 * self.N], np.int32)
        bv = np.ascontiguousarray(bondv, np.float32)
        bn = np.ascontiguousarray(bondn, np.int32)
        self._tbc = (bt, bv, bn)
        check(lib().i3g_markers_set_tbc(self.h, _p(bt), _p(bv), _p(bn), C.c_long(len(bn))), "markers_set_tbc")

    def update_tk(self, tk, tk2, kt, ro, cp, heatdif, timestep, zdeep=1e30, tdeep=0.0):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (tk, tk2, kt, ro, cp)]
        check(lib().i3g_markers_update_tk(self.h, *[_p(a) for a in arrs], C.c_double(heatdif),
                                          C.c_double(timestep), C.c_double(zdeep), C.c_double(tdeep)),
              "markers_update_tk")

    def advect_rk4(self, dt, vx, vy, vz):
        v = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
        check(lib().i3g_markers_advect_rk4(self.h, C.c_double(dt), *[_p(a) for a in v]), "markers_rk4")

    def advect_ex(self, dt, vx, vy, vz, outgrid=0, markmod=0, vyfluid=0.0, vymelt=0.0, g=(0.0, 9.81, 0.0), pr=None,
                  markd=None):
        """movemark's motion (i3g_markers_advect_ex); markd by marker id"""
        v = [np.ascontiguousarray(a, np.float32) for a in (vx, vy, vz)]
        mp = MoveParams(int(outgrid), int(markmod), float(vyfluid), float(vymelt), (C.c_double * 3)(*g))
        prd = None if pr is None else np.ascontiguousarray(pr, np.float64)
        md = None if markd is None else np.ascontiguousarray(markd, np.float32)
        check(lib().i3g_markers_advect_ex(self.h, C.c_double(dt), *[_p(a) for a in v], C.byref(mp), _p(prd), _p(md)),
              "markers_advect_ex")

    def set_grain(self, rr, rn):
        a, b = (np.ascontiguousarray(q, np.float32) for q in (rr, rn))
        check(lib().i3g_markers_set_grain(self.h, _p(a), _p(b)), "markers_set_grain")

    def set_order(self, order):
        """node averages: 1 cell-segmented (default), 0 the reference's
        serial marker order"""
        check(lib().i3g_markers_set_order(self.h, C.c_int(order)), "markers_set_order")

    def set_melt(self, w, ex):
        """water content and extracted melt by marker id"""
        a, b = (np.ascontiguousarray(q, np.float32) for q in (w, ex))
        check(lib().i3g_markers_set_melt(self.h, C.c_long(len(a)), _p(a), _p(b)), "markers_set_melt")

    def get_grain(self):
        rr, rn = np.zeros(self.n, np.float32), np.zeros(self.n, np.float32)
        check(lib().i3g_markers_get_grain(self.h, _p(rr), _p(rn)), "markers_get_grain")
        return rr, rn

    def set_stress(self, st12):
        if st12 is None:
            check(lib().i3g_markers_set_stress(self.h, None), "markers_set_stress")
            return
        self._st = [np.ascontiguousarray(a, np.float32) for a in st12]
        sp = (C.c_void_p * 12)(*[a.ctypes.data for a in self._st])
        check(lib().i3g_markers_set_stress(self.h, sp), "markers_set_stress")


# ---- heat -------------------------------------------------------------------

class HeatParams(C.Structure):
    _fields_ = [("timestep", C.c_double), ("maxtkstep", C.c_double), ("heatmin", C.c_double),
                ("t0koef", C.c_double), ("sweeps", C.c_int), ("max_sweeps", C.c_int),
                ("timesum", C.c_double), ("timebeg", C.c_double), ("timeend", C.c_double),
                ("tempend", C.c_double), ("verbose", C.c_int)]


class HeatStats(C.Structure):
    _fields_ = [("converged", C.c_int), ("substeps", C.c_int), ("sweeps", C.c_int),
                ("heatsum", C.c_double), ("bondsum", C.c_double), ("seconds", C.c_double),
                ("device_seconds", C.c_double)]


class Heat:
    """Device heat context for one level-0 grid (i3g_heat_create .. destroy)."""

    def __init__(self, gx, gy, gz, bondm, bondv, bondn):
        self._keep = [np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz)] + [
            np.ascontiguousarray(bondm, np.int32), np.ascontiguousarray(bondv, np.float32),
            np.ascontiguousarray(bondn, np.int32)]
        gx, gy, gz, bm, bv, bn = self._keep
        self.N = len(gx) * len(gy) * len(gz)
        s = Setup(len(gx), len(gy), len(gz), _p(gx), _p(gy), _p(gz), 0, _p(bm), _p(bv), _p(bn), len(bn))
        self.h = C.c_void_p()
        check(lib().i3g_heat_create(C.byref(s), C.byref(self.h)), "heat_create")

    def close(self):
        if self.h:
            lib().i3g_heat_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sources(self, adiabyn=0, frictyn=0, gk=(0.0, 9.81, 0.0), ysize=0.0, stp100=0.0, vx=None, vy=None,
                    vz=None, pr=None, et=None, sig=None, eps=None):
        """Adiabatic (adiabyn 1/2) and shear (frictyn) heating terms of the
        following solves (i3g_heat_set_sources); sig/eps: 6 planes each."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        keep = [f(vx), f(vy), f(vz), None if pr is None else np.ascontiguousarray(pr, np.float64), f(et)]
        s6 = [f(a) for a in sig] if sig is not None else None
        e6 = [f(a) for a in eps] if eps is not None else None
        sp = (C.c_void_p * 6)(*[a.ctypes.data for a in s6]) if s6 else None
        ep = (C.c_void_p * 6)(*[a.ctypes.data for a in e6]) if e6 else None
        gkd = np.ascontiguousarray(gk, np.float64)
        check(lib().i3g_heat_set_sources(self.h, C.c_int(adiabyn), C.c_int(frictyn), _p(gkd), C.c_double(ysize),
                                         C.c_double(stp100), *[_p(a) for a in keep], sp, ep), "heat_set_sources")

    def solve(self, tk, ro, cp, kt, ht, check_conv=True, **kw):
        """Returns (new tk float32, HeatStats); tk is not modified."""
        p = HeatParams()
        lib().i3g_heat_default_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        out = np.array(tk, np.float32, copy=True)
        fl = [np.ascontiguousarray(a, np.float32) for a in (ro, cp, kt, ht)]
        st = HeatStats()
        rc = lib().i3g_heat_solve(self.h, _p(out), *[_p(a) for a in fl], C.byref(p), C.byref(st))
        if rc != -4 or check_conv:
            check(rc, "heat_solve")
        return out, st


# ---- learned warm start -----------------------------------------------------

class WarmStartNet:
    """Device encoder-decoder (i3g_warmstart_*) on an (nz, nx, ny) node grid."""

    def __init__(self, nz, nx, ny, seed=0):
        self.shape = (nz, nx, ny)
        self.N = nz * nx * ny
        self.h = C.c_void_p()
        check(lib().i3g_warmstart_create(C.c_long(nz), C.c_long(nx), C.c_long(ny), C.c_ulonglong(seed),
                                         C.byref(self.h)), "warmstart_create")
        lib().i3g_warmstart_nparams.restype = C.c_long
        self.nparams = lib().i3g_warmstart_nparams(self.h)

    def close(self):
        if self.h:
            lib().i3g_warmstart_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        check(lib().i3g_warmstart_get_params(self.h, _p(p)), "warmstart_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib().i3g_warmstart_set_params(self.h, _p(p)), "warmstart_set_params")

    def infer(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(self.N, 2)
        y = np.zeros((self.N, 4), np.float32)
        check(lib().i3g_warmstart_infer(self.h, _p(x), _p(y)), "warmstart_infer")
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(self.N, 2)
        t = np.ascontiguousarray(t, np.float32).reshape(self.N, 4)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        check(lib().i3g_warmstart_grad(self.h, _p(x), _p(t), C.byref(loss), _p(g)), "warmstart_grad")
        return loss.value, g

    def train_step(self, xs, ts, lr=1e-3):
        xs = np.ascontiguousarray(xs, np.float32)
        ts = np.ascontiguousarray(ts, np.float32)
        b = xs.size // (2 * self.N)
        loss = C.c_double()
        check(lib().i3g_warmstart_train_step(self.h, C.c_int(b), _p(xs), _p(ts), C.c_double(lr), C.byref(loss)),
              "warmstart_train_step")
        return loss.value


WARMSTART_LAYERS = [(2, 16, 1), (16, 32, 2), (32, 64, 2), (64, 64, 2), (64, 32, 1), (32, 16, 1), (16, 4, 1)]


def unflatten_params(p):
    """device parameter vector -> [W0, b0, W1, b1, ...] with W [27, cout, cin]"""
    out, o = [], 0
    for ci, co, _ in WARMSTART_LAYERS:
        out.append(np.asarray(p[o:o + 27 * co * ci]).reshape(27, co, ci)); o += 27 * co * ci
        out.append(np.asarray(p[o:o + co])); o += co
    return out


# ---- the time-step loop ------------------------------------------------------

def make_markers(cfg, per_axis, seed=4):
    """seeded markers of configuration cfg['which'] (i3g_gen_markers):
    x, y, z, tk (float32) and type (int8)"""
    args = (C.c_int(cfg["which"]), C.c_long(cfg["nx"]), C.c_long(cfg["ny"]), C.c_long(cfg["nz"]),
            C.c_ulonglong(seed), _p(cfg["gx"]), _p(cfg["gy"]), _p(cfg["gz"]), C.c_int(per_axis))
    lib().i3g_gen_markers.restype = C.c_long
    n = check(lib().i3g_gen_markers(*args, None, None, None, None, None), "gen_markers")
    out = [np.zeros(n, np.float32) for _ in range(4)] + [np.zeros(n, np.int8)]
    check(lib().i3g_gen_markers(*args, *[_p(a) for a in out]), "gen_markers")
    return out


def arrhenius_rocks(eta_ref, t_ref=1600.0, e_act=3.0e5, n0=1e18, n1=1e25, ro=None, bro=3e-5, aro=1e-3, cp=1000.0,
                    kt=3.0, ht=0.0):
    """Rock table (types 0..len-1) with Arrhenius viscosity
    eta = A exp(E / (R T)) (viscalc branch B, mark3mg.c:1278-1299), A chosen
    so that eta(t_ref) = eta_ref[type], capped to [n0, n1]"""
    rocks = []
    for t, e in enumerate(eta_ref):
        a = e / np.exp(e_act / (8.314 * t_ref)) if e > 0 else 0.0
        rocks.append(dict(nu0=a, dh=e_act if e > 0 else 0.0, dv=0.0, n0=n0, n1=n1,
                          ro=(ro[t] if ro is not None else 3300.0), bro=bro, aro=aro, cp=cp, kt=kt, kf=0.0, kp=0.0,
                          ht=ht))
    return rocks


class ModelParams(C.Structure):
    _fields_ = [("mk", MkParams), ("stokes", StokesParams), ("heat", HeatParams), ("maxtmstep", C.c_double),
                ("maxxystep", C.c_double), ("heatdif", C.c_double), ("pscale", C.c_double), ("vscale", C.c_double),
                ("learned", C.c_int)]


class ModelStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("t_stokes", C.c_double), ("t_heat", C.c_double), ("t_mktk", C.c_double),
                ("t_rk4", C.c_double), ("t_sort", C.c_double), ("t_transfer", C.c_double), ("vcycles", C.c_int),
                ("outer", C.c_int), ("rel_res", C.c_double), ("timestep", C.c_double), ("timesum", C.c_double),
                ("learned_taken", C.c_int), ("heat_substeps", C.c_int), ("heat_sweeps", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def model_params(rtol=1e-8, maxtmstep=1e12, maxxystep=None, heatdif=0.0, nubeg=1e18, nuend=1e25, densimod=1,
                 heatmin=1e-6, learned=False, pscale=1.0, vscale=1.0, **stokes_kw):
    """ModelParams with the library's defaults for the solve and the heat
    step"""
    p = ModelParams()
    p.mk.nubeg, p.mk.nuend, p.mk.nukoef, p.mk.densimod = nubeg, nuend, 0.0, densimod
    p.mk.timestep, p.mk.vdeep, p.mk.nudeep, p.mk.viscmod = 0.0, 1e30, 1e30, 0
    lib().i3g_stokes_default_params(C.byref(p.stokes))
    p.stokes.rtol = rtol
    for k, v in stokes_kw.items():
        setattr(p.stokes, k, v)
    lib().i3g_heat_default_params(C.byref(p.heat))
    p.heat.heatmin = heatmin
    p.maxtmstep = maxtmstep
    p.maxxystep = maxxystep if maxxystep is not None else 0.0
    p.heatdif = heatdif
    p.pscale, p.vscale, p.learned = pscale, vscale, 1 if learned else 0
    return p


class Model:
    """The resident time-step loop (i3g_model_*): vpiterate, titerate,
    movemark, ronurecalc per step on one GPU."""

    def __init__(self, cfg, capacity):
        self.cfg = cfg
        self.N = cfg["nx"] * cfg["ny"] * cfg["nz"]
        self._keep = [np.ascontiguousarray(cfg[k]) for k in ("gx", "gy", "gz", "bondm", "bondv", "bondn")]
        gx, gy, gz, bm, bv, bn = self._keep
        s = Setup(cfg["nx"], cfg["ny"], cfg["nz"], _p(gx), _p(gy), _p(gz), cfg["multinum"],
                  _p(bm), _p(bv), _p(bn), len(bn))
        self.h = C.c_void_p()
        check(lib().i3g_model_create(C.byref(s), C.c_long(capacity), C.byref(self.h)), "model_create")
        phys = np.array([cfg["g"][0], cfg["g"][1], cfg["g"][2], cfg["nubeg"], cfg["nuend"], cfg["viscmod"]])
        check(lib().i3g_model_set_physics(self.h, _p(phys)), "model_set_physics")

    def close(self):
        if self.h:
            lib().i3g_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_markers(self, x, y, z, tk, mtype, rocks):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, tk)]
        mt = np.ascontiguousarray(mtype, np.int8)
        ra = rocks_array(rocks)
        check(lib().i3g_model_set_markers(self.h, C.c_long(len(arrs[0])), *[_p(a) for a in arrs], _p(mt), ra,
                                          C.c_int(len(rocks))), "model_set_markers")

    def init(self, params, pr=None):
        pp = None if pr is None else _p(np.ascontiguousarray(pr, np.float64))
        check(lib().i3g_model_init(self.h, C.byref(params), pp), "model_init")

    def step(self, params, net=None):
        st = ModelStats()
        check(lib().i3g_model_step(self.h, C.byref(params), net.h if net is not None else None, C.byref(st)),
              "model_step")
        return st

    def get(self):
        N = self.N
        out = dict(vx=np.zeros(N, np.float32), vy=np.zeros(N, np.float32), vz=np.zeros(N, np.float32),
                   pr=np.zeros(N), tk=np.zeros(N, np.float32), nu=np.zeros(N, np.float32))
        check(lib().i3g_model_get(self.h, *[_p(out[k]) for k in ("vx", "vy", "vz", "pr", "tk", "nu")]), "model_get")
        return out

    def stokes_ctx(self):
        lib().i3g_model_stokes.restype = C.c_void_p
        return lib().i3g_model_stokes(self.h)

    def markers_handle(self):
        lib().i3g_model_markers.restype = C.c_void_p
        return lib().i3g_model_markers(self.h)


# ---- run files of the reference (host only) ---------------------------------

_MODE_D1 = ("tkpor", "zmpor", "vyfluid", "vymelt", "dmwamin", "tdeep", "dtdeep", "drdeep", "zdeep", "vdeep",
            "nudeep")
_MODE_D2 = ("dxwater", "dywater", "dzwater", "dserp", "yserp", "tserp", "lambmlt", "ydike", "dydike", "nusseltno")
_MODE_D3 = ("circdepth", "circtemp", "healrate", "meltmin", "maxwater", "maxpmelt", "teclmin", "teclmax", "xvolc",
            "xtkplut")


class ModeC(C.Structure):
    _fields_ = ([("fl0cap", C.c_int), ("fl0out", C.c_void_p), ("fl0otp", C.c_void_p), ("fl0cyc", C.c_void_p),
                 ("fl0stp", C.c_void_p), ("fln3", C.c_int), ("fl1in", C.c_char * 64), ("fl1itp", C.c_int),
                 ("fl0num", C.c_int)]
                + [(n, C.c_int) for n in ("printmod", "movemod", "tempmod", "markmod", "gridmod", "outgrid",
                                          "densimod", "filesjob")]
                + [(n, C.c_double) for n in ("stp100", "eroslev", "sedilev", "waterlev")]
                + [("cyc1max", C.c_int), ("multinum", C.c_int), ("multicyc", C.c_int),
                   ("multinnn", (C.c_int * MAXLEV) * 2)]
                + [(n, C.c_double) for n in ("DIVVMIN", "STOKSMIN", "DIVVMAX", "STOKSMAX")]
                + [(n, C.c_double) for n in ("p0koef", "p1koef", "p2koef", "v0koef", "v1koef", "nubeg", "nuend",
                                             "nukoef")]
                + [("viscmod", C.c_int), ("spheryn", C.c_int)]
                + [("cyc2max", C.c_int), ("multinumt", C.c_int), ("multicyct", C.c_int),
                   ("multittt", (C.c_int * MAXLEV) * 2)]
                + [(n, C.c_double) for n in ("HEATMIN", "t0koef", "t1koef", "heatdif")]
                + [("frictyn", C.c_int), ("adiabyn", C.c_int)]
                + [(n, C.c_double) for n in _MODE_D1 + _MODE_D2 + _MODE_D3]
                + [(n, C.c_double) for n in ("timebeg", "timeend", "tempbeg", "convend")])


def read_mode(mode_path, file_path=None, cap=5000):
    """The parameter deck (i3g_mode_read): a dict of the reference's
    loadconf variables plus the output-file table"""
    m = ModeC()
    out = np.zeros((cap, 64), np.uint8)
    otp = np.zeros(cap, np.int32)
    cyc = np.zeros((cap, 4), np.int32)
    stp = np.zeros((cap, 11))
    m.fl0cap = cap
    m.fl0out, m.fl0otp, m.fl0cyc, m.fl0stp = (a.ctypes.data for a in (out, otp, cyc, stp))
    check(lib().i3g_mode_read(str(mode_path).encode(), None if file_path is None else str(file_path).encode(),
                              C.byref(m)), "mode_read")
    k = m.fl0num
    d = {f: getattr(m, f) for f, _ in ModeC._fields_ if f not in ("fl0cap", "fl0out", "fl0otp", "fl0cyc", "fl0stp",
                                                                   "multinnn", "multittt", "fl1in")}
    d["fl1in"] = m.fl1in.decode()
    d["fl0out"] = [bytes(out[i]).split(b"\0", 1)[0].decode() for i in range(k)]
    d["fl0otp"], d["fl0cyc"], d["fl0stp"] = otp[:k].copy(), cyc[:k].copy(), stp[:k].copy()
    d["multinnn"] = np.array([[m.multinnn[a][b] for b in range(m.multinum + 1)] for a in range(2)])
    d["multittt"] = np.array([[m.multittt[a][b] for b in range(m.multinumt + 1)] for a in range(2)])
    return d


class PrnC(C.Structure):
    _fields_ = ([(n, C.c_long) for n in ("xnumx", "ynumy", "znumz", "mnumx", "mnumy", "mnumz")]
                + [(n, C.c_double) for n in ("xsize", "ysize", "zsize")]
                + [(n, C.c_long) for n in ("pxinit", "pyinit", "pzinit")]
                + [(n, C.c_double) for n in ("pinit", "gxkoef", "gykoef", "gzkoef")]
                + [("rocknum", C.c_int), ("bondnum", C.c_long), ("marknum", C.c_long), ("n0", C.c_int)]
                + [(n, C.c_double) for n in ("timesum_years", "gridcur", "gridtot")]
                + [(n, C.c_void_p) for n in ("rock", "node", "gx", "gy", "gz", "bond_at", "bond_rec",
                                             "bond_partner", "bond_v", "mark", "markt")])


PRN_ARRAYS = ("rock", "node", "gx", "gy", "gz", "bond_at", "bond_rec", "bond_partner", "bond_v", "mark", "markt")


def read_prn(path):
    """The .prn state file (i3g_prn_read): header scalars and arrays (rock
    24 x rocknum, node 16 x nodenum, mark 12 x marknum, bond_v bondnum x 2)"""
    h = PrnC()
    check(lib().i3g_prn_read_header(str(path).encode(), C.byref(h)), "prn_read_header")
    N = h.xnumx * h.ynumy * h.znumz
    a = dict(rock=np.zeros((24, h.rocknum)), node=np.zeros((16, N), np.float32), gx=np.zeros(h.xnumx, np.float32),
             gy=np.zeros(h.ynumy, np.float32), gz=np.zeros(h.znumz, np.float32),
             bond_at=np.zeros(h.bondnum, np.int64), bond_rec=np.zeros(h.bondnum, np.int64),
             bond_partner=np.zeros(h.bondnum, np.int64), bond_v=np.zeros((h.bondnum, 2), np.float32),
             mark=np.zeros((12, h.marknum), np.float32), markt=np.zeros(h.marknum, np.int8))
    for k in PRN_ARRAYS:
        setattr(h, k, a[k].ctypes.data if a[k].size else None)
    check(lib().i3g_prn_read(str(path).encode(), C.byref(h)), "prn_read")
    d = {f: getattr(h, f) for f, _ in PrnC._fields_ if f not in PRN_ARRAYS}
    d.update(a)
    return d


def write_prn(path, d):
    """i3g_prn_write from a dict shaped as read_prn returns"""
    h = PrnC()
    for f, _ in PrnC._fields_:
        if f not in PRN_ARRAYS:
            setattr(h, f, d[f])
    dt = dict(rock=np.float64, node=np.float32, gx=np.float32, gy=np.float32, gz=np.float32, bond_at=np.int64,
              bond_rec=np.int64, bond_partner=np.int64, bond_v=np.float32, mark=np.float32, markt=np.int8)
    keep = {k: np.ascontiguousarray(d[k], dt[k]) for k in PRN_ARRAYS}
    for k in PRN_ARRAYS:
        setattr(h, k, keep[k].ctypes.data if keep[k].size else None)
    check(lib().i3g_prn_write(str(path).encode(), C.byref(h)), "prn_write")


# ---- initial condition generator and ParaView output -----------------------


class RockBox(C.Structure):
    """i3g_rock_box"""
    _fields_ = [("shape", C.c_int), ("type", C.c_int), ("ex", C.c_double), ("e", C.c_double),
                ("x", C.c_double * 8), ("y", C.c_double * 8), ("z", C.c_double * 8)]


class TempBox(C.Structure):
    """i3g_temp_box"""
    _fields_ = ([("type", C.c_int)] +
                [(f, C.c_double) for f in ("x0", "x2", "y0", "y1", "y2", "y3", "z0", "x5", "x7", "y4", "y5", "y6",
                                           "y7", "z7")] + [("t", C.c_double * 8)])


def deck_boxes(path, xsize, ysize, zsize, maxbox=4096):
    """rock and temperature boxes of a deck in the reference's format
    (i3g_deck_boxes): (RockBox array, TempBox array)"""
    rb = (RockBox * maxbox)()
    tb = (TempBox * maxbox)()
    nr, nt = C.c_int(), C.c_int()
    check(lib().i3g_deck_boxes(str(path).encode(), C.c_double(xsize), C.c_double(ysize), C.c_double(zsize), rb,
                               C.c_int(maxbox), C.byref(nr), tb, C.c_int(maxbox), C.byref(nt)), "deck_boxes")
    return (RockBox * nr.value)(*rb[:nr.value]), (TempBox * nt.value)(*tb[:nt.value])


def _paint(self, xsize, ysize, zsize, boxes, ex=None, od=None, ow=None):
    """i3g_markers_paint; ex/od/ow (current marker order) updated copies"""
    n = self.n
    outs = [None if a is None else np.array(a, np.float32, copy=True) for a in (ex, od, ow)]
    for a in outs:
        if a is not None and a.size != n:
            raise ValueError("marker arrays of the wrong length")
    check(lib().i3g_markers_paint(self.h, C.c_double(xsize), C.c_double(ysize), C.c_double(zsize), boxes,
                                  C.c_int(len(boxes)), *[_p(a) for a in outs]), "markers_paint")
    return outs


Markers.paint = _paint


def temperature_boxes(gx, gy, gz, xsize, ysize, zsize, boxes, tk):
    """i3g_temperature_boxes on a copy of the node temperatures"""
    g = [np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz)]
    out = np.array(tk, np.float32, copy=True)
    check(lib().i3g_temperature_boxes(C.c_long(len(g[0])), C.c_long(len(g[1])), C.c_long(len(g[2])),
                                      *[_p(a) for a in g], C.c_double(xsize), C.c_double(ysize), C.c_double(zsize),
                                      boxes, C.c_int(len(boxes)), _p(out)), "temperature_boxes")
    return out


def write_vtr_nodal(path, gx, gy, gz, fields):
    """fields: {name: node plane (m3, m1, m2 order)}; float64 planes are
    written as Float64, others as Float32"""
    g = [np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz)]
    names = list(fields)
    f64 = [np.asarray(fields[k]).dtype == np.float64 for k in names]
    arrs = [np.ascontiguousarray(fields[k], np.float64 if d else np.float32) for k, d in zip(names, f64)]
    cn = (C.c_char_p * len(names))(*[k.encode() for k in names])
    cf = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    cd = (C.c_int * len(names))(*[int(d) for d in f64])
    check(lib().i3g_write_vtr_nodal(str(path).encode(), C.c_long(len(g[0])), C.c_long(len(g[1])),
                                    C.c_long(len(g[2])), *[_p(a) for a in g], C.c_int(len(names)), cn, cf, cd),
          "write_vtr_nodal")


def write_vtr_composition(path, xsize, ysize, zsize, mx, my, mz, x, y, z, mtype):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    t = np.ascontiguousarray(mtype, np.int8)
    check(lib().i3g_write_vtr_composition(str(path).encode(), C.c_double(xsize), C.c_double(ysize),
                                          C.c_double(zsize), C.c_long(mx), C.c_long(my), C.c_long(mz),
                                          C.c_long(len(t)), *[_p(a) for a in arrs], _p(t)), "write_vtr_composition")


def pvd_open(path):
    check(lib().i3g_pvd_open(str(path).encode()), "pvd_open")


def pvd_append(path, time, datafile, directory=None):
    check(lib().i3g_pvd_append(str(path).encode(), C.c_double(time), datafile.encode(),
                               None if directory is None else directory.encode()), "pvd_append")
