"""Oracle — TEST INFRASTRUCTURE ONLY.

A plain-C FP64 restatement of the reference Stokes operator, transfers and
norms (``oracle/i3g_oracle.c``), loaded through ctypes.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may use it,
and only as the checker: the product path never calls into this package.

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
    x = np.array(x, np.float64, copy=True)
    lib().or_gs_sweep(C.byref(L), *[_p(e) for e in etas], _p(bondm), _p(bondv),
                      _p(bondn), _p(b), _p(x), C.c_double(p0koef), C.c_double(v0koef))
    return x
