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
                    level-0 system (configuration 0 geometry) by an
                    equilibrated sparse direct solve, pressure gauge
                    P(2,1,2) = 0 as vpiterate (move3mg.c:2248-2269)
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
    x = D * spl.spsolve(As.tocsc(), Rs * b)
    r = rhs - (A @ x)
    r[ref] = 0
    np.savez_compressed(os.path.join(HERE, f"g3_converged_c{which}_{n}.npz"), n=n, which=which, seed=seed,
                        sol=x, rhs=rhs, exists=ex, resid_max=np.abs(r).max())
    print(f"g3_converged_c{which}_{n}.npz written; max |b - A x| = {np.abs(r).max():.3e}")


if __name__ == "__main__":
    R = Ref()
    R.lib.refh_set_threads(1)
    g1(R)
    g3(R, 17)
    g3(R, 25)
