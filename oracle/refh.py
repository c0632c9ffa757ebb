"""ctypes view of the reference harness oracle/_ref/libi3ref.so — TEST
INFRASTRUCTURE ONLY.

The harness is the reference's own solver sources compiled from the
reference checkout (see oracle/Makefile); this module sets the reference's
global state and calls its routines.  Used to write golden vectors under
tests/golden/ (in the container that holds the reference) and to time the
reference CPU path for the benchmark's CPU baseline.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, "_ref", "libi3ref.so")
# the same reference unit plus the product's drop-in glue (oracle/Makefile: glue)
GLUE_PATH = os.path.join(_HERE, "_ref", "libi3ref_glue.so")


def available(path=PATH):
    return os.path.isfile(path)


class Ref:
    def __init__(self, path=PATH):
        self.lib = C.CDLL(path)
        caps = (C.c_long * 11)()
        self.lib.refh_caps(caps)
        (self.MAXXLN, self.MAXYLN, self.MAXZLN, self.MAXNOD, self.MAXPPP,
         self.MAXPAR, self.MAXBON, self.MAXMRK, self.MAXNNN, self.MAXMAT,
         self.MAXTMR) = list(caps)
        self.lib.refh_vpiterate.restype = C.c_double
        for f in ("refh_m1serch", "refh_m2serch", "refh_m3serch"):
            getattr(self.lib, f).restype = C.c_long
            getattr(self.lib, f).argtypes = [C.c_double]
        self.lib.refh_assemble_level.restype = C.c_long

    # -- raw global views -------------------------------------------------
    def arr(self, name, ctype, n):
        return np.ctypeslib.as_array((ctype * n).in_dll(self.lib, name))

    def f32(self, name, n=None):
        return self.arr(name, C.c_float, n or self.MAXNOD)

    def f64(self, name, n=None):
        return self.arr(name, C.c_double, n or self.MAXNOD)

    def i32(self, name, n):
        return self.arr(name, C.c_int, n)

    def set(self, name, value, ctype=C.c_double):
        ctype.in_dll(self.lib, name).value = value

    def get(self, name, ctype=C.c_double):
        return ctype.in_dll(self.lib, name).value

    # -- setup ------------------------------------------------------------
    def setup_grid(self, gx, gy, gz, multinum):
        gx, gy, gz = (np.ascontiguousarray(a, np.float64) for a in (gx, gy, gz))
        rc = self.lib.refh_setup_grid(C.c_long(len(gx)), C.c_long(len(gy)),
                                      C.c_long(len(gz)), gx.ctypes.data_as(C.c_void_p),
                                      gy.ctypes.data_as(C.c_void_p),
                                      gz.ctypes.data_as(C.c_void_p), C.c_int(multinum))
        if rc:
            raise ValueError("grid too large for the harness capacities")
        self.nlev = multinum + 1
        self.mgx = self.i32("mgx", 100)[:self.nlev].copy()
        self.mgy = self.i32("mgy", 100)[:self.nlev].copy()
        self.mgz = self.i32("mgz", 100)[:self.nlev].copy()
        self.mgp = self.i32("mgp", 100)[:self.nlev].copy()
        self.mgn = self.i32("mgn", 100)[:self.nlev].copy()
        self.nodenum = len(gx) * len(gy) * len(gz)

    def level_lines(self, l):
        gx = self.arr("mggx", C.c_double, 10 * self.MAXXLN).reshape(10, self.MAXXLN)
        gy = self.arr("mggy", C.c_double, 10 * self.MAXYLN).reshape(10, self.MAXYLN)
        gz = self.arr("mggz", C.c_double, 10 * self.MAXZLN).reshape(10, self.MAXZLN)
        return (gx[l, :self.mgx[l]].copy(), gy[l, :self.mgy[l]].copy(),
                gz[l, :self.mgz[l]].copy())

    def set_bc(self, bondm, bondv, bondn):
        self.i32("bondm", self.MAXPAR)[:len(bondm)] = bondm
        bv = self.arr("bondv1", C.c_float, 2 * self.MAXBON)
        bv[:len(bondv)] = bondv
        self.i32("bondn1", self.MAXBON)[:len(bondn)] = bondn

    def set_physics(self, g=(0.0, 9.81, 0.0), nubeg=1e18, nuend=1e25, viscmod=0,
                    p0koef=0.3, p1koef=0.1, p2koef=0.0, v0koef=0.7, v1koef=0.7):
        self.set("GXKOEF", g[0]); self.set("GYKOEF", g[1]); self.set("GZKOEF", g[2])
        self.set("nubeg", nubeg); self.set("nuend", nuend)
        self.set("viscmod", viscmod, C.c_int)
        for k, v in dict(p0koef=p0koef, p1koef=p1koef, p2koef=p2koef,
                         v0koef=v0koef, v1koef=v1koef).items():
            self.set(k, v)

    def assemble_level(self, l):
        """CSR rows of level l as the reference stores them (global
        column indices), plus the right-hand side."""
        n = int(self.mgn[l]) * 4
        rowptr = np.zeros(n + 1, np.int64)
        cols = np.zeros(n * 17, np.int64)
        vals = np.zeros(n * 17)
        rhs = np.zeros(n)
        nnz = self.lib.refh_assemble_level(C.c_int(l), rowptr.ctypes.data_as(C.c_void_p),
                                           cols.ctypes.data_as(C.c_void_p),
                                           vals.ctypes.data_as(C.c_void_p),
                                           rhs.ctypes.data_as(C.c_void_p))
        return rowptr, cols[:nnz], vals[:nnz], rhs
