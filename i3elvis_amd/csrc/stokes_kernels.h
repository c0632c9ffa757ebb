/*
 * Device kernels of the i3g Stokes path (gfx950).  Templated on the real
 * type T (double for the FP64 residual/plain multigrid, float for the
 * mixed-precision V-cycle).  Included by i3g_stokes.hip only.
 *
 * Layout per level: node n = m3*nx*ny + m1*ny + m2; an unknown vector is
 * 4 SoA planes of N (P, vx, vy, vz).  flags[n]: bit c = unknown c is a
 * boundary-condition row, bit 4+c = unknown c exists on the staggered grid.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FL_BC(c) (1u << (c))
#define FL_EX(c) (1u << (4 + (c)))

/* Per-level geometry factors (device arrays, T precision):
 *   ih[i] = 1/(g[i+1]-g[i])   (i < n-1)
 *   ic[i] = 1/(g[i+1]-g[i-1]) (0 < i < n-1)
 * as used by sxxcalc (cell width), sxycalc (two-cell span) and
 * x/y/zstokserr (move3mg.c:16-18, 369-370, 729-731). */
template <typename T> struct LevelView {
	int nx, ny, nz;
	long N;
	const T *ihx, *ihy, *ihz, *icx, *icy, *icz;
	const T *eC, *eXY, *eXZ, *eYZ;
	const uint8_t *flags;
	const T *wP; /* continuity-row weight eta_cell / mean cell size (P plane) */
};

template <typename T> struct Vec4 {
	T *p, *x, *y, *z;
};
template <typename T> struct CVec4 {
	const T *p, *x, *y, *z;
};

/* ---- stencil: A x and diagonal of one row (x/y/zstokserr, continerr) --- */

template <typename T, typename X>
__device__ __forceinline__ T sxx_v(const LevelView<T> &L, const X *vx, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	long v0 = n - sz - sx - 1;
	return (T)2 * L.eC[n] * ((T)vx[v0 + sx] - (T)vx[v0]) * L.ihx[a - 1];
}
template <typename T, typename X>
__device__ __forceinline__ T syy_v(const LevelView<T> &L, const X *vy, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	long v0 = n - sz - sx - 1;
	return (T)2 * L.eC[n] * ((T)vy[v0 + 1] - (T)vy[v0]) * L.ihy[b - 1];
}
template <typename T, typename X>
__device__ __forceinline__ T szz_v(const LevelView<T> &L, const X *vz, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	long v0 = n - sz - sx - 1;
	return (T)2 * L.eC[n] * ((T)vz[v0 + sz] - (T)vz[v0]) * L.ihz[c - 1];
}
template <typename T, typename X>
__device__ __forceinline__ T sxy_v(const LevelView<T> &L, const X *vx, const X *vy, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	return (T)2 * L.eXY[n] * (((T)vx[n] - (T)vx[n - 1]) * L.icy[b] + ((T)vy[n] - (T)vy[n - sx]) * L.icx[a]);
}
template <typename T, typename X>
__device__ __forceinline__ T sxz_v(const LevelView<T> &L, const X *vx, const X *vz, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	return (T)2 * L.eXZ[n] * (((T)vx[n] - (T)vx[n - sz]) * L.icz[c] + ((T)vz[n] - (T)vz[n - sx]) * L.icx[a]);
}
template <typename T, typename X>
__device__ __forceinline__ T syz_v(const LevelView<T> &L, const X *vy, const X *vz, int a, int b, int c)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = c * sz + a * sx + b;
	return (T)2 * L.eYZ[n] * (((T)vy[n] - (T)vy[n - sz]) * L.icz[c] + ((T)vz[n] - (T)vz[n - 1]) * L.icy[b]);
}

/* A x for the stencil row of component comp at (i,j,k); diag optional. */
template <typename T, typename X>
__device__ __forceinline__ T row_ax(const LevelView<T> &L, const X *P, const X *vx,
				    const X *vy, const X *vz, int comp, int i, int j,
				    int k, T *diag)
{
	long sx = L.ny, sz = (long)L.nx * L.ny;
	long n = k * sz + i * sx + j;
	T ax, d;
	if (comp == 1) {
		T xkf = (T)2 * L.icx[i], ykf = L.ihy[j], zkf = L.ihz[k];
		ax = xkf * (sxx_v(L, vx, i + 1, j + 1, k + 1) - sxx_v(L, vx, i, j + 1, k + 1)) +
		     ykf * (sxy_v(L, vx, vy, i, j + 1, k) - sxy_v(L, vx, vy, i, j, k)) +
		     zkf * (sxz_v(L, vx, vz, i, j, k + 1) - sxz_v(L, vx, vz, i, j, k)) -
		     xkf * ((T)P[n + sz + sx + 1] - (T)P[n + sz + 1]);
		if (diag)
			*diag = -xkf * ((T)2 * L.eC[n + sz + sx + 1] * L.ihx[i] + (T)2 * L.eC[n + sz + 1] * L.ihx[i - 1]) -
				ykf * ((T)2 * L.eXY[n + 1] * L.icy[j + 1] + (T)2 * L.eXY[n] * L.icy[j]) -
				zkf * ((T)2 * L.eXZ[n + sz] * L.icz[k + 1] + (T)2 * L.eXZ[n] * L.icz[k]);
	} else if (comp == 2) {
		T xkf = L.ihx[i], ykf = (T)2 * L.icy[j], zkf = L.ihz[k];
		ax = ykf * (syy_v(L, vy, i + 1, j + 1, k + 1) - syy_v(L, vy, i + 1, j, k + 1)) +
		     xkf * (sxy_v(L, vx, vy, i + 1, j, k) - sxy_v(L, vx, vy, i, j, k)) +
		     zkf * (syz_v(L, vy, vz, i, j, k + 1) - syz_v(L, vy, vz, i, j, k)) -
		     ykf * ((T)P[n + sz + sx + 1] - (T)P[n + sz + sx]);
		if (diag)
			*diag = -ykf * ((T)2 * L.eC[n + sz + sx + 1] * L.ihy[j] + (T)2 * L.eC[n + sz + sx] * L.ihy[j - 1]) -
				xkf * ((T)2 * L.eXY[n + sx] * L.icx[i + 1] + (T)2 * L.eXY[n] * L.icx[i]) -
				zkf * ((T)2 * L.eYZ[n + sz] * L.icz[k + 1] + (T)2 * L.eYZ[n] * L.icz[k]);
	} else if (comp == 3) {
		T xkf = L.ihx[i], ykf = L.ihy[j], zkf = (T)2 * L.icz[k];
		ax = zkf * (szz_v(L, vz, i + 1, j + 1, k + 1) - szz_v(L, vz, i + 1, j + 1, k)) +
		     xkf * (sxz_v(L, vx, vz, i + 1, j, k) - sxz_v(L, vx, vz, i, j, k)) +
		     ykf * (syz_v(L, vy, vz, i, j + 1, k) - syz_v(L, vy, vz, i, j, k)) -
		     zkf * ((T)P[n + sz + sx + 1] - (T)P[n + sx + 1]);
		if (diag)
			*diag = -zkf * ((T)2 * L.eC[n + sz + sx + 1] * L.ihz[k] + (T)2 * L.eC[n + sx + 1] * L.ihz[k - 1]) -
				xkf * ((T)2 * L.eXZ[n + sx] * L.icx[i + 1] + (T)2 * L.eXZ[n] * L.icx[i]) -
				ykf * ((T)2 * L.eYZ[n + 1] * L.icy[j + 1] + (T)2 * L.eYZ[n] * L.icy[j]);
	} else {
		long v0 = n - sz - sx - 1;
		ax = ((T)vx[v0 + sx] - (T)vx[v0]) * L.ihx[i - 1] +
		     ((T)vy[v0 + 1] - (T)vy[v0]) * L.ihy[j - 1] +
		     ((T)vz[v0 + sz] - (T)vz[v0]) * L.ihz[k - 1];
		d = 0;
		if (diag) *diag = d;
	}
	return ax;
}

/* Boundary-condition record of one unknown (bonderr, move3mg.c:1532-1579):
 * x_u = c + k * x_partner.  partner = plane*N + node, or -1. */
template <typename T> struct BcRec {
	const int *u;        /* plane*N + node */
	const int *partner;  /* plane*N + node or -1 */
	const T *c, *k;
};

/* ------------------------------------------------------------------ */
/* Kernels                                                            */
/* ------------------------------------------------------------------ */

#define RBLK 256

template <typename T>
__device__ __forceinline__ void block_sum2(double a, double b, double *out2)
{
	__shared__ double sa[RBLK], sb[RBLK];
	int t = threadIdx.x;
	sa[t] = a; sb[t] = b;
	__syncthreads();
	for (int s = RBLK / 2; s > 0; s >>= 1) {
		if (t < s) { sa[t] += sa[t + s]; sb[t] += sb[t + s]; }
		__syncthreads();
	}
	if (t == 0) { out2[0] = sa[0]; out2[1] = sb[0]; }
}

/* r = b - A x (mode 0) or r = W A x (mode 1: FGMRES operator, P plane
 * scaled by wP).  Thread per node: all four unknowns.  BC rows:
 * r = c - (x_u - k x_partner).  When partials != NULL, each block writes the
 * sums of (w r)^2 and (w b)^2 over non-BC rows (continuity rows weighted
 * by wP) for the convergence metric. Block = 256 threads over the xy plane,
 * grid.y = z. */
template <typename T, typename X, int MODE>
__device__ __forceinline__ void residual_node(const LevelView<T> &L, const X *__restrict__ x,
					      const T *__restrict__ b, T *__restrict__ r,
					      const int *__restrict__ bcid, const BcRec<T> &bc, int i, int j, int k,
					      bool sums, double *sr, double *sb)
{
	long N = L.N;
	long n = (long)k * L.nx * L.ny + (long)i * L.ny + j;
	unsigned fl = L.flags[n];
	const X *P = x, *vx = x + N, *vy = x + 2 * N, *vz = x + 3 * N;
#pragma unroll
	for (int c = 0; c < 4; c++) {
		T res = 0;
		long u = c * N + n;
		if (fl & FL_EX(c)) {
			if (fl & FL_BC(c)) {
				int rec = bcid[u];
				int p = bc.partner[rec];
				T ax = (T)x[u] - (p >= 0 ? bc.k[rec] * (T)x[p] : (T)0);
				res = MODE ? ax : (b ? b[u] : (T)0) - ax;
			} else {
				T ax = row_ax<T, X>(L, P, vx, vy, vz, c, i, j, k, (T *)0);
				T bb = b ? b[u] : (T)0;
				res = MODE ? ax : bb - ax;
				if (MODE && c == 0) res *= L.wP[n];
				if (sums) {
					double w = c == 0 ? (double)L.wP[n] : 1.0;
					*sr += ((double)res * w) * ((double)res * w);
					*sb += ((double)bb * w) * ((double)bb * w);
				}
			}
		}
		r[u] = res;
	}
}

template <typename T, typename X, int MODE>
__global__ void __launch_bounds__(256)
k_residual(LevelView<T> L, const X *__restrict__ x, const T *__restrict__ b,
	   T *__restrict__ r, const int *__restrict__ bcid, BcRec<T> bc,
	   double *partials)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int k = blockIdx.y;
	int i = t / L.ny, j = t - i * L.ny;
	double sr = 0, sb = 0;
	if (i < L.nx) residual_node<T, X, MODE>(L, x, b, r, bcid, bc, i, j, k, partials != nullptr, &sr, &sb);
	if (partials) {
		long bid = (long)blockIdx.y * gridDim.x + blockIdx.x;
		block_sum2<T>(sr, sb, partials + 2 * bid);
	}
}

/* Gauss-Seidel relaxation of velocity component comp on the unknowns of one
 * colour ((i+j+k) mod 2 == colour): x += v0koef (b - A x)/diag, the
 * gausmat4 update (move3mg.c:3121-3124).  BC rows are left to k_bc.
 * Threads cover every other y index of each (x,z) line. */
/* Work-group -> (plane, chunk) map.  remap 0: grid (chunks, planes).
 * remap 1: a 1D grid of 8*ceil(chunks*planes/8) groups; the hardware hands
 * groups to the 8 XCDs round-robin, so group g is given work item
 * (g % 8) * per + g / 8: every XCD sweeps a contiguous slab of planes in
 * order and finds the neighbouring planes in its own L2. */
__device__ __forceinline__ bool block_work(int remap, int chunks, int planes, int k0, int *chunk, int *plane)
{
	if (!remap) {
		*chunk = blockIdx.x;
		*plane = k0 + (int)blockIdx.y;
		return true;
	}
	int per = gridDim.x >> 3;
	int w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
	if (w >= chunks * planes) return false;
	int q = w / chunks;
	*chunk = w - q * chunks;
	*plane = k0 + q;
	return true;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_smooth_vel(LevelView<T> L, T *__restrict__ x, const T *__restrict__ b, int comp,
	     int colour, T v0koef, int remap, int chunks, int k0, int planes)
{
	int ch, k;
	if (!block_work(remap, chunks, planes, k0, &ch, &k)) return;
	int t = ch * blockDim.x + threadIdx.x;
	int hy = (L.ny + 1) >> 1;
	int i = t / hy, tj = t - i * hy;
	if (i >= L.nx) return;
	int j = 2 * tj + ((i + k + colour) & 1);
	if (j >= L.ny) return;
	long N = L.N;
	long n = (long)k * L.nx * L.ny + (long)i * L.ny + j;
	unsigned fl = L.flags[n];
	if (!(fl & FL_EX(comp)) || (fl & FL_BC(comp))) return;
	T d;
	T ax = row_ax<T, T>(L, x, x + N, x + 2 * N, x + 3 * N, comp, i, j, k, &d);
	long u = comp * N + n;
	T res = (b ? b[u] : (T)0) - ax;
	x[u] += v0koef * res / d;
}

/* The same relaxation with every operand of the row loaded up front: the
 * plane-per-launch kernel above is latency-bound (its loads are consumed
 * as they arrive, about eight dependent memory round trips per wave); here
 * the flag is read first and then all 25 loads of the row (the unknowns
 * of the stencil, the stress-point viscosities, the rhs) are issued before
 * any is used.  Same arithmetic as row_ax; the scheduling barrier keeps the
 * compiler from sinking the loads back to their uses.  COMP = 1, 2, 3.
 *
 * vel_relax_a returns the relaxed value x + v0koef (b - A x)/diag of
 * component COMP at one node (gausmat4, move3mg.c:3121-3124).  Every operand
 * comes through the accessor `a`, at offsets (di, dj, dk) from the node:
 * a.u<c>() the unknowns (0 P, 1 vx, 2 vy, 3 vz), a.eC()/eXY()/eXZ()/eYZ()
 * the stress-point viscosities, a.ihx(d)/icx(d) (and y, z) the line
 * factors.  Every smoother kernel, whatever memory it reads from, calls
 * this one function, so that they all round alike. */
template <typename T, int COMP, class A>
__device__ __forceinline__ T vel_relax_a(const A &a, T rhs, T v0koef)
{
	T ax, d, a0;
	if (COMP == 1) {
		T c0 = a.eC(1, 1, 1), c1 = a.eC(0, 1, 1);
		T xy1 = a.eXY(0, 1, 0), xy0 = a.eXY(0, 0, 0), xz1 = a.eXZ(0, 0, 1), xz0 = a.eXZ(0, 0, 0);
		a0 = a.template u<1>(0, 0, 0);
		T axp = a.template u<1>(1, 0, 0), axm = a.template u<1>(-1, 0, 0);
		T ayp = a.template u<1>(0, 1, 0), aym = a.template u<1>(0, -1, 0);
		T azp = a.template u<1>(0, 0, 1), azm = a.template u<1>(0, 0, -1);
		T by1 = a.template u<2>(0, 1, 0), by1x = a.template u<2>(-1, 1, 0);
		T by0 = a.template u<2>(0, 0, 0), by0x = a.template u<2>(-1, 0, 0);
		T cz1 = a.template u<3>(0, 0, 1), cz1x = a.template u<3>(-1, 0, 1);
		T cz0 = a.template u<3>(0, 0, 0), cz0x = a.template u<3>(-1, 0, 0);
		T p1 = a.template u<0>(1, 1, 1), p0 = a.template u<0>(0, 1, 1);
		T ihx1 = a.ihx(0), ihx0 = a.ihx(-1), icx = a.icx(0), icy1 = a.icy(1), icy0 = a.icy(0);
		T icz1 = a.icz(1), icz0 = a.icz(0), ihy = a.ihy(0), ihz = a.ihz(0);
		__builtin_amdgcn_sched_barrier(0);
		T xkf = (T)2 * icx, ykf = ihy, zkf = ihz;
		T sxx1 = (T)2 * c0 * (axp - a0) * ihx1, sxx0 = (T)2 * c1 * (a0 - axm) * ihx0;
		T sxy1 = (T)2 * xy1 * ((ayp - a0) * icy1 + (by1 - by1x) * icx);
		T sxy0 = (T)2 * xy0 * ((a0 - aym) * icy0 + (by0 - by0x) * icx);
		T sxz1 = (T)2 * xz1 * ((azp - a0) * icz1 + (cz1 - cz1x) * icx);
		T sxz0 = (T)2 * xz0 * ((a0 - azm) * icz0 + (cz0 - cz0x) * icx);
		ax = xkf * (sxx1 - sxx0) + ykf * (sxy1 - sxy0) + zkf * (sxz1 - sxz0) - xkf * (p1 - p0);
		d = -xkf * ((T)2 * c0 * ihx1 + (T)2 * c1 * ihx0) - ykf * ((T)2 * xy1 * icy1 + (T)2 * xy0 * icy0) -
		    zkf * ((T)2 * xz1 * icz1 + (T)2 * xz0 * icz0);
	} else if (COMP == 2) {
		T c0 = a.eC(1, 1, 1), c1 = a.eC(1, 0, 1);
		T xy1 = a.eXY(1, 0, 0), xy0 = a.eXY(0, 0, 0), yz1 = a.eYZ(0, 0, 1), yz0 = a.eYZ(0, 0, 0);
		a0 = a.template u<2>(0, 0, 0);
		T ayp = a.template u<2>(0, 1, 0), aym = a.template u<2>(0, -1, 0);
		T axp = a.template u<2>(1, 0, 0), axm = a.template u<2>(-1, 0, 0);
		T azp = a.template u<2>(0, 0, 1), azm = a.template u<2>(0, 0, -1);
		T bx1 = a.template u<1>(1, 0, 0), bx1y = a.template u<1>(1, -1, 0);
		T bx0 = a.template u<1>(0, 0, 0), bx0y = a.template u<1>(0, -1, 0);
		T cz1 = a.template u<3>(0, 0, 1), cz1y = a.template u<3>(0, -1, 1);
		T cz0 = a.template u<3>(0, 0, 0), cz0y = a.template u<3>(0, -1, 0);
		T p1 = a.template u<0>(1, 1, 1), p0 = a.template u<0>(1, 0, 1);
		T ihy1 = a.ihy(0), ihy0 = a.ihy(-1), icy = a.icy(0), icx1 = a.icx(1), icx0 = a.icx(0);
		T icz1 = a.icz(1), icz0 = a.icz(0), ihx = a.ihx(0), ihz = a.ihz(0);
		__builtin_amdgcn_sched_barrier(0);
		T xkf = ihx, ykf = (T)2 * icy, zkf = ihz;
		T syy1 = (T)2 * c0 * (ayp - a0) * ihy1, syy0 = (T)2 * c1 * (a0 - aym) * ihy0;
		T sxy1 = (T)2 * xy1 * ((bx1 - bx1y) * icy + (axp - a0) * icx1);
		T sxy0 = (T)2 * xy0 * ((bx0 - bx0y) * icy + (a0 - axm) * icx0);
		T syz1 = (T)2 * yz1 * ((azp - a0) * icz1 + (cz1 - cz1y) * icy);
		T syz0 = (T)2 * yz0 * ((a0 - azm) * icz0 + (cz0 - cz0y) * icy);
		ax = ykf * (syy1 - syy0) + xkf * (sxy1 - sxy0) + zkf * (syz1 - syz0) - ykf * (p1 - p0);
		d = -ykf * ((T)2 * c0 * ihy1 + (T)2 * c1 * ihy0) - xkf * ((T)2 * xy1 * icx1 + (T)2 * xy0 * icx0) -
		    zkf * ((T)2 * yz1 * icz1 + (T)2 * yz0 * icz0);
	} else {
		T c0 = a.eC(1, 1, 1), c1 = a.eC(1, 1, 0);
		T xz1 = a.eXZ(1, 0, 0), xz0 = a.eXZ(0, 0, 0), yz1 = a.eYZ(0, 1, 0), yz0 = a.eYZ(0, 0, 0);
		a0 = a.template u<3>(0, 0, 0);
		T azp = a.template u<3>(0, 0, 1), azm = a.template u<3>(0, 0, -1);
		T axp = a.template u<3>(1, 0, 0), axm = a.template u<3>(-1, 0, 0);
		T ayp = a.template u<3>(0, 1, 0), aym = a.template u<3>(0, -1, 0);
		T bx1 = a.template u<1>(1, 0, 0), bx1z = a.template u<1>(1, 0, -1);
		T bx0 = a.template u<1>(0, 0, 0), bx0z = a.template u<1>(0, 0, -1);
		T cy1 = a.template u<2>(0, 1, 0), cy1z = a.template u<2>(0, 1, -1);
		T cy0 = a.template u<2>(0, 0, 0), cy0z = a.template u<2>(0, 0, -1);
		T p1 = a.template u<0>(1, 1, 1), p0 = a.template u<0>(1, 1, 0);
		T ihz1 = a.ihz(0), ihz0 = a.ihz(-1), icz = a.icz(0), icx1 = a.icx(1), icx0 = a.icx(0);
		T icy1 = a.icy(1), icy0 = a.icy(0), ihx = a.ihx(0), ihy = a.ihy(0);
		__builtin_amdgcn_sched_barrier(0);
		T xkf = ihx, ykf = ihy, zkf = (T)2 * icz;
		T szz1 = (T)2 * c0 * (azp - a0) * ihz1, szz0 = (T)2 * c1 * (a0 - azm) * ihz0;
		T sxz1 = (T)2 * xz1 * ((bx1 - bx1z) * icz + (axp - a0) * icx1);
		T sxz0 = (T)2 * xz0 * ((bx0 - bx0z) * icz + (a0 - axm) * icx0);
		T syz1 = (T)2 * yz1 * ((cy1 - cy1z) * icz + (ayp - a0) * icy1);
		T syz0 = (T)2 * yz0 * ((cy0 - cy0z) * icz + (a0 - aym) * icy0);
		ax = zkf * (szz1 - szz0) + xkf * (sxz1 - sxz0) + ykf * (syz1 - syz0) - zkf * (p1 - p0);
		d = -zkf * ((T)2 * c0 * ihz1 + (T)2 * c1 * ihz0) - xkf * ((T)2 * xz1 * icx1 + (T)2 * xz0 * icx0) -
		    ykf * ((T)2 * yz1 * icy1 + (T)2 * yz0 * icy0);
	}
	return a0 + v0koef * (rhs - ax) / d;
}

/* Operands of a row in global memory: the unknown planes P, vx, vy, vz and
 * the level's viscosity and line-factor arrays at node n = (i, j, k), except
 * unknown OC, which comes through the functor `f(di, dj, dk)`. */
template <typename T> struct OwnGlobal {
	const T *a;
	long n, sx, sz;
	__device__ __forceinline__ T operator()(int di, int dj, int dk) const { return a[n + di * sx + dj + dk * sz]; }
};

template <typename T, int OC, class F> struct AccGlobal {
	const LevelView<T> &L;
	const T *P, *vx, *vy, *vz;
	long n, sx, sz;
	int i, j, k;
	const F &f;
	__device__ __forceinline__ long o(int di, int dj, int dk) const { return n + di * sx + dj + dk * sz; }
	template <int C> __device__ __forceinline__ T u(int di, int dj, int dk) const
	{
		if (C == OC) return f(di, dj, dk);
		const T *p = C == 0 ? P : (C == 1 ? vx : (C == 2 ? vy : vz));
		return p[o(di, dj, dk)];
	}
	__device__ __forceinline__ T eC(int di, int dj, int dk) const { return L.eC[o(di, dj, dk)]; }
	__device__ __forceinline__ T eXY(int di, int dj, int dk) const { return L.eXY[o(di, dj, dk)]; }
	__device__ __forceinline__ T eXZ(int di, int dj, int dk) const { return L.eXZ[o(di, dj, dk)]; }
	__device__ __forceinline__ T eYZ(int di, int dj, int dk) const { return L.eYZ[o(di, dj, dk)]; }
	__device__ __forceinline__ T ihx(int d) const { return L.ihx[i + d]; }
	__device__ __forceinline__ T icx(int d) const { return L.icx[i + d]; }
	__device__ __forceinline__ T ihy(int d) const { return L.ihy[j + d]; }
	__device__ __forceinline__ T icy(int d) const { return L.icy[j + d]; }
	__device__ __forceinline__ T ihz(int d) const { return L.ihz[k + d]; }
	__device__ __forceinline__ T icz(int d) const { return L.icz[k + d]; }
};

/* vel_relax_a with the component's own seven stencil values through
 * `own(di, dj, dk)` and everything else from global memory */
template <typename T, int COMP, class Own>
__device__ __forceinline__ T vel_relax(const LevelView<T> &L, const T *__restrict__ P, const T *__restrict__ vx,
				       const T *__restrict__ vy, const T *__restrict__ vz, T rhs, int i, int j, int k,
				       long n, T v0koef, const Own &own)
{
	AccGlobal<T, COMP, Own> a{L, P, vx, vy, vz, n, (long)L.ny, (long)L.nx * L.ny, i, j, k, own};
	return vel_relax_a<T, COMP>(a, rhs, v0koef);
}

template <typename T, int COMP>
__device__ __forceinline__ void vel_update(const LevelView<T> &L, T *__restrict__ x, const T *__restrict__ b,
					   int i, int j, int k, T v0koef)
{
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	long n = (long)k * sz + (long)i * sx + j;
	unsigned fl = L.flags[n];
	if (!(fl & FL_EX(COMP)) || (fl & FL_BC(COMP))) return;
	long u = COMP * N + n;
	T rhs = b ? b[u] : (T)0;
	OwnGlobal<T> own{x + COMP * N, n, sx, sz};
	x[u] = vel_relax<T, COMP>(L, x, x + N, x + 2 * N, x + 3 * N, rhs, i, j, k, n, v0koef, own);
}

template <typename T, int COMP>
__global__ void __launch_bounds__(256)
k_smooth_vel_ilp(LevelView<T> L, T *__restrict__ x, const T *__restrict__ b, int colour, T v0koef, int remap,
		 int chunks, int k0, int planes)
{
	int ch, k;
	if (!block_work(remap, chunks, planes, k0, &ch, &k)) return;
	int t = ch * blockDim.x + threadIdx.x;
	int hy = (L.ny + 1) >> 1;
	int i = t / hy, tj = t - i * hy;
	if (i >= L.nx) return;
	int j = 2 * tj + ((i + k + colour) & 1);
	if (j >= L.ny) return;
	vel_update<T, COMP>(L, x, b, i, j, k, v0koef);
}

/* Pressure relaxation P -= p0koef * eta_cell * (div v - b_P)
 * (move3mg.c:3105-3111); BC rows are left to k_bc.  p_relax_a returns the
 * relaxed value from the old p, the operands through an accessor as in
 * vel_relax_a; p_relax reads vz through `vzf(di, dj, dk)`. */
template <typename T, class A>
__device__ __forceinline__ T p_relax_a(const A &a, T p, T rhs, T p0koef)
{
	T ax = (a.template u<1>(0, -1, -1) - a.template u<1>(-1, -1, -1)) * a.ihx(-1) +
	       (a.template u<2>(-1, 0, -1) - a.template u<2>(-1, -1, -1)) * a.ihy(-1) +
	       (a.template u<3>(-1, -1, 0) - a.template u<3>(-1, -1, -1)) * a.ihz(-1);
	T res = ax - rhs;
	return p - res * a.eC(0, 0, 0) * p0koef;
}

template <typename T, class VZ>
__device__ __forceinline__ T p_relax(const LevelView<T> &L, const T *__restrict__ vx, const T *__restrict__ vy,
				     const VZ &vzf, T p, T rhs, int i, int j, int k, long n, T p0koef)
{
	AccGlobal<T, 3, VZ> a{L, (const T *)nullptr, vx, vy, (const T *)nullptr, n, (long)L.ny, (long)L.nx * L.ny, i, j, k, vzf};
	return p_relax_a<T>(a, p, rhs, p0koef);
}

template <typename T>
__device__ __forceinline__ void p_update(const LevelView<T> &L, T *__restrict__ x, const T *__restrict__ b, int i,
					 int j, int k, T p0koef)
{
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	long n = (long)k * sz + (long)i * sx + j;
	unsigned fl = L.flags[n];
	if (!(fl & FL_EX(0)) || (fl & FL_BC(0))) return;
	OwnGlobal<T> vzf{x + 3 * N, n, sx, sz};
	x[n] = p_relax<T>(L, x + N, x + 2 * N, vzf, x[n], b ? b[n] : (T)0, i, j, k, n, p0koef);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_smooth_p(LevelView<T> L, T *__restrict__ x, const T *__restrict__ b, T p0koef, int k0)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int k = k0 + (int)blockIdx.y;
	int i = t / L.ny, j = t - i * L.ny;
	if (i >= L.nx) return;
	p_update<T>(L, x, b, i, j, k, p0koef);
}

/* ---- both colours of one velocity component in one pass ----------------- */

/* The colour-ordered sweep (vx colour 0, vx colour 1, vy 0, vy 1, vz 0, vz 1,
 * P) as three launches, one per velocity component, the pressure update
 * folded into the vz launch.  Each launch reads the sweep's input vector
 * and writes a second vector (out of place), so every old value a work-group
 * needs stays readable while the others write.
 *
 * A work-group owns an (x, y) tile of a z-slab and marches up through the
 * slab plane by plane.  The component's old values of the tile plus two
 * rings are staged in an LDS ring of four planes.  At plane kk it relaxes
 * colour 0 on the tile plus one ring (the ring's values are the neighbouring
 * work-group's, recomputed here from the same old inputs, bit for bit),
 * then colour 1 of plane kk-1 on the tile from the colour-0 values in LDS,
 * then stores plane kk-1 of the tile, and in the vz launch relaxes P of
 * plane kk-1 (the tile shifted by +1 in x and y: P of cell (i,j,k) reads vz
 * at (i-1,j-1,k) and (i-1,j-1,k-1)).  The vz launch relaxes colour 1 one
 * plane past its slab (not stored) so that P of the slab's top plane sees
 * the new vz.  Every node update goes through vel_relax / p_relax, so the
 * result equals the seven phase launches of the in-place sweep to the bit.
 * Unknowns that are not relaxed (boundary rows, missing unknowns) are
 * copied to the output; P of the planes i = 0, j = 0, k = 0 (no unknowns)
 * is never written and stays zero. */
struct FusedGeom {
	int ntx, nty, nzs; /* tiles in x, y; slabs in z */
	int txm, tym;      /* largest tile extents */
	int remap;         /* XCD-contiguous work order */
};

template <typename T> struct OwnLds {
	const T *m, *z, *p; /* slots of planes k-1, k, k+1 at node (i,j) */
	int rs;             /* row stride (y) */
	__device__ __forceinline__ T operator()(int di, int dj, int dk) const
	{
		const T *s = dk < 0 ? m : (dk > 0 ? p : z);
		return s[di * rs + dj];
	}
};

template <typename T, int COMP>
__global__ void __launch_bounds__(256)
k_sweep_comp(LevelView<T> L, const T *__restrict__ P, const T *__restrict__ vx, const T *__restrict__ vy,
	     const T *__restrict__ vz, const T *__restrict__ b, T *__restrict__ out, T *__restrict__ pout,
	     FusedGeom G, T v0koef, T p0koef)
{
	extern __shared__ __align__(16) unsigned char lds_raw[];
	T *ring = (T *)lds_raw;
	int w, dummy;
	int total = G.ntx * G.nty * G.nzs;
	if (!block_work(G.remap, total, 1, 0, &w, &dummy)) return;
	int tx = w % G.ntx, rest = w / G.ntx;
	int ty = rest % G.nty, sl = rest / G.nty;
	const int nx = L.nx, ny = L.ny, nz = L.nz;
	const long N = L.N, sx = ny, sz = (long)nx * ny;
	const int i0 = (int)((long)tx * nx / G.ntx), i1 = (int)((long)(tx + 1) * nx / G.ntx);
	const int j0 = (int)((long)ty * ny / G.nty), j1 = (int)((long)(ty + 1) * ny / G.nty);
	const int k0 = (int)((long)sl * nz / G.nzs), k1 = (int)((long)(sl + 1) * nz / G.nzs);
	const int rs = G.tym + 4, slot = (G.txm + 4) * rs;
	const T *own_old = COMP == 1 ? vx : (COMP == 2 ? vy : vz);
	const T *bown = b ? b + COMP * N : nullptr;
	/* region bounds: load (tile + 2 rings), colour 0 (tile + 1 ring) */
	const int la = max(i0 - 2, 0), lb = min(i1 + 2, nx), lc = max(j0 - 2, 0), ld = min(j1 + 2, ny);
	const int ca = max(i0 - 1, 0), cb = min(i1 + 1, nx), cc = max(j0 - 1, 0), cd = min(j1 + 1, ny);
	const int c1end = COMP == 3 ? min(k1 + 1, nz) : k1;
	auto lidx = [&](int i, int j) { return (i - i0 + 2) * rs + (j - j0 + 2); };
	auto load_plane = [&](int q) {
		T *s = ring + (q & 3) * slot;
		int wl = ld - lc, cnt = (lb - la) * wl;
		long base = (long)q * sz;
		for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
			int ii = t / wl, jj = t - ii * wl;
			int i = la + ii, j = lc + jj;
			s[lidx(i, j)] = own_old[base + (long)i * sx + j];
		}
	};
	/* relax colour `col` of plane kq on [ia,ib) x [ja,jb) */
	auto relax = [&](int kq, int col, int ia, int ib, int ja, int jb) {
		T *sz0 = ring + (kq & 3) * slot;
		const T *sm = ring + ((kq + 3) & 3) * slot, *sp = ring + ((kq + 1) & 3) * slot;
		int hw = (jb - ja + 1) >> 1, cnt = (ib - ia) * hw;
		for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
			int ii = t / hw, jj = t - ii * hw;
			int i = ia + ii;
			int j = ja + ((col - i - kq - ja) & 1) + 2 * jj;
			if (j >= jb) continue;
			long n = (long)kq * sz + (long)i * sx + j;
			unsigned fl = L.flags[n];
			if (!(fl & FL_EX(COMP)) || (fl & FL_BC(COMP))) continue;
			T rhs = bown ? bown[n] : (T)0;
			int li = lidx(i, j);
			OwnLds<T> own{sm + li, sz0 + li, sp + li, rs};
			sz0[li] = vel_relax<T, COMP>(L, P, vx, vy, vz, rhs, i, j, kq, n, v0koef, own);
		}
	};
	for (int q = k0 - 2; q < k0; q++)
		if (q >= 0) load_plane(q);
	for (int kk = k0 - 1; kk <= c1end; kk++) {
		if (kk + 1 < nz) load_plane(kk + 1);
		__syncthreads();
		if (kk >= 0 && kk < nz) relax(kk, 0, ca, cb, cc, cd);
		__syncthreads();
		int kc = kk - 1;
		if (kc >= k0 && kc < c1end) relax(kc, 1, i0, i1, j0, j1);
		__syncthreads();
		if (kc >= k0 && kc < k1) {
			const T *s = ring + (kc & 3) * slot;
			int wt = j1 - j0, cnt = (i1 - i0) * wt;
			long base = (long)kc * sz;
			for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
				int ii = t / wt, jj = t - ii * wt;
				int i = i0 + ii, j = j0 + jj;
				out[base + (long)i * sx + j] = s[lidx(i, j)];
			}
		}
		if (COMP == 3 && kc >= k0 + 1 && kc <= k1 && kc >= 1 && kc < nz) {
			const T *s0 = ring + (kc & 3) * slot, *s1 = ring + ((kc + 3) & 3) * slot;
			int pa = max(i0 + 1, 1), pb = min(i1 + 1, nx), pc = max(j0 + 1, 1), pd = min(j1 + 1, ny);
			int wt = pd - pc, cnt = (pb - pa) * wt;
			for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
				int ii = t / wt, jj = t - ii * wt;
				int i = pa + ii, j = pc + jj;
				long n = (long)kc * sz + (long)i * sx + j;
				unsigned fl = L.flags[n];
				T pold = P[n];
				if ((fl & FL_EX(0)) && !(fl & FL_BC(0))) {
					int li = lidx(i, j);
					OwnLds<T> vzf{s1 + li, s0 + li, s0 + li, rs};
					pold = p_relax<T>(L, vx, vy, vzf, pold, b ? b[n] : (T)0, i, j, kc, n, p0koef);
				}
				pout[n] = pold;
			}
		}
		__syncthreads();
	}
}

/* ---- the whole sweep in one launch: a z-march through LDS ---------------- */

/* One FP32 sweep (phases vx c0, vx c1, vy c0, vy c1, vz c0, vz c1, P) in one
 * launch, out of place (xin -> xout), equal to the seven phase launches of
 * the in-place sweep to the bit.
 *
 * How far behind the march front a phase can run.  A row reads one plane up
 * and down, but a phase needs the previous phase's output one plane up only
 * where it relaxes the second colour of a component (the first colour at
 * k+1); the other phases read the earlier components at their own plane and
 * below (vy rows read vx at k; vz rows vx, vy at k and k-1; P rows the
 * velocities at k and k-1).  Conversely a phase must still find old values
 * where it reads a later phase's unknowns: vx/vy rows read vz and P at k+1,
 * vz rows P at k and k+1, and every first colour the component's second
 * colour at k-1.  Phase q therefore runs SW7_LAG(q) = 0,1,1,2,2,3,3 planes
 * behind phase 0 (phases of one step in order 0..6): every node is relaxed
 * from the values the phase-by-phase sweep gives it.
 *
 * A work-group owns an (x, y) tile of at most SW7_T x SW7_T nodes and
 * marches up through all planes, two per step: at step kk (even) phase q
 * relaxes planes kk - lag and kk - lag + 1.  Phase q relaxes the tile plus
 * the lines that the later phases of the tile read (its region R[q],
 * clamped to the grid): the halo values are the neighbouring work-groups'
 * own, computed by the same function from the same inputs.  The four
 * unknowns and four viscosities of the staged region (tile + 5 lines on
 * both sides in x, 6 below and 5 above in y) are held in LDS in one ring of
 * planes per field, deep enough for the planes the step reads plus the two
 * it loads for the next step; the next step's planes, right-hand sides and
 * flags are loaded into registers while the current step computes.  Phase 6
 * writes the tile's four unknowns of its two planes to xout (every node of
 * the grid exactly once). */
#define SW7_T 17
#define SW7_RX (SW7_T + 10)
#define SW7_RY (SW7_T + 11)
#define SW7_PL (SW7_RX * SW7_RY)
#define SW7_THREADS 1024
#define SW7_SLOTS 47
#define SW7_MAXZ 1024
#ifndef SW7_LOAD_PHASE
#define SW7_LOAD_PHASE 2 /* phase before which the next step's planes are loaded */
#endif

/* LDS fields: 0 P, 1 vx, 2 vy, 3 vz, 4 eC, 5 eXY, 6 eXZ, 7 eYZ.  Ring depth
 * = planes read below the step's base plane kk (dn) + planes read above
 * kk+1 (up) + 2; sw7_base: first slot of each ring. */
__host__ __device__ constexpr int sw7_up(int f) { return (f == 0 || f == 1 || f == 3 || f == 4 || f == 6) ? 1 : 0; }
__host__ __device__ constexpr int sw7_depth(int f)
{
	return f == 0 ? 6 : f == 1 ? 7 : f == 2 ? 6 : f == 3 ? 7 : f == 4 ? 6 : f == 5 ? 4 : f == 6 ? 6 : 5;
}
__host__ __device__ constexpr int sw7_base(int f)
{
	return f == 0 ? 0 : f == 1 ? 6 : f == 2 ? 13 : f == 3 ? 19 : f == 4 ? 26 : f == 5 ? 32 : f == 6 ? 36 : 42;
}
/* float offset of field f's slot holding plane pl (pl >= -8) */
__device__ __forceinline__ int sw7_slot(int f, int pl) { return (sw7_base(f) + (pl + 8) % sw7_depth(f)) * SW7_PL; }

/* phase q: lag, component, colour, region margins (x below, x above, y below, y above) */
__host__ __device__ constexpr int sw7_lag(int q) { return q == 0 ? 0 : q <= 2 ? 1 : q <= 4 ? 2 : 3; }
__host__ __device__ constexpr int sw7_comp(int q) { return q == 6 ? 0 : 1 + q / 2; }
__host__ __device__ constexpr int sw7_mx0(int q) { return q == 0 ? 4 : q <= 2 ? 3 : q <= 4 ? 2 : q == 5 ? 1 : 0; }
__host__ __device__ constexpr int sw7_mx1(int q) { return q == 0 ? 4 : q == 1 ? 3 : q == 2 ? 2 : q <= 4 ? 1 : 0; }
__host__ __device__ constexpr int sw7_my0(int q) { return q == 0 ? 5 : q == 1 ? 4 : q == 2 ? 3 : q <= 4 ? 2 : q == 5 ? 1 : 0; }
__host__ __device__ constexpr int sw7_my1(int q) { return q == 0 ? 4 : q <= 2 ? 3 : q == 3 ? 2 : q == 4 ? 1 : 0; }

struct Sw7Acc {
	const float *s;          /* LDS rings */
	int o, li, lj, pl;       /* node: offset in a region plane, region line x, y; plane */
	const float *gxh, *gxc, *gyh, *gyc, *gzh, *gzc;
	__device__ __forceinline__ float at(int f, int di, int dj, int dk) const
	{
		return s[sw7_slot(f, pl + dk) + o + di * SW7_RY + dj];
	}
	template <int C> __device__ __forceinline__ float u(int di, int dj, int dk) const { return at(C, di, dj, dk); }
	__device__ __forceinline__ float eC(int di, int dj, int dk) const { return at(4, di, dj, dk); }
	__device__ __forceinline__ float eXY(int di, int dj, int dk) const { return at(5, di, dj, dk); }
	__device__ __forceinline__ float eXZ(int di, int dj, int dk) const { return at(6, di, dj, dk); }
	__device__ __forceinline__ float eYZ(int di, int dj, int dk) const { return at(7, di, dj, dk); }
	__device__ __forceinline__ float ihx(int d) const { return gxh[li + d]; }
	__device__ __forceinline__ float icx(int d) const { return gxc[li + d]; }
	__device__ __forceinline__ float ihy(int d) const { return gyh[lj + d]; }
	__device__ __forceinline__ float icy(int d) const { return gyc[lj + d]; }
	__device__ __forceinline__ float ihz(int d) const { return gzh[pl + d]; }
	__device__ __forceinline__ float icz(int d) const { return gzc[pl + d]; }
};

/* Work-group barrier for LDS only: waits for this wave's LDS operations,
 * not for its global loads in flight (__syncthreads() also drains those,
 * which would expose the next step's prefetch at every phase). */
__device__ __forceinline__ void sw7_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

struct Sw7Geom {
	int ntx, nty;   /* tiles in x and y */
	int remap;      /* XCD-contiguous tile order */
};

/* dynamic LDS bytes of k_sweep7 for a level of nz planes */
static inline size_t sw7_lds_bytes(int nz)
{
	return sizeof(float) * ((size_t)SW7_SLOTS * SW7_PL + 2 * SW7_RX + 2 * SW7_RY + 2 * (size_t)nz);
}

__global__ void __launch_bounds__(SW7_THREADS)
k_sweep7(LevelView<float> L, const float *__restrict__ xin, float *__restrict__ xout, const float *__restrict__ b,
	 Sw7Geom G, float v0koef, float p0koef)
{
	extern __shared__ __align__(16) unsigned char lds_raw[];
	float *s = (float *)lds_raw;
	float *gxh = s + SW7_SLOTS * SW7_PL, *gxc = gxh + SW7_RX, *gyh = gxc + SW7_RX, *gyc = gyh + SW7_RY;
	float *gzh = gyc + SW7_RY, *gzc = gzh + L.nz;
	int w, dummy;
	if (!block_work(G.remap, G.ntx * G.nty, 1, 0, &w, &dummy)) return;
	const int tx = w / G.nty, ty = w - tx * G.nty;
	const int nx = L.nx, ny = L.ny, nz = L.nz;
	const long N = L.N, sz = (long)nx * ny;
	const int i0 = (int)((long)tx * nx / G.ntx), i1 = (int)((long)(tx + 1) * nx / G.ntx);
	const int j0 = (int)((long)ty * ny / G.nty), j1 = (int)((long)(ty + 1) * ny / G.nty);
	const int rx0 = i0 - 5, ry0 = j0 - 6; /* region origin (may lie outside the grid) */
	const int t = threadIdx.x;
	const int pidx = t >> 9, r = t & 511; /* half of the group: plane kk-lag+pidx */

	for (int q = t; q < SW7_RX; q += SW7_THREADS) {
		int i = rx0 + q;
		bool in = i >= 0 && i < nx;
		gxh[q] = in ? L.ihx[i] : 0.f;
		gxc[q] = in ? L.icx[i] : 0.f;
	}
	for (int q = t; q < SW7_RY; q += SW7_THREADS) {
		int j = ry0 + q;
		bool in = j >= 0 && j < ny;
		gyh[q] = in ? L.ihy[j] : 0.f;
		gyc[q] = in ? L.icy[j] : 0.f;
	}
	for (int q = t; q < nz; q += SW7_THREADS) {
		gzh[q] = L.ihz[q];
		gzc[q] = L.icz[q];
	}

	/* this thread's node of each phase (same every step): region offsets
	 * li << 8 | lj, or -1 */
	int code[7];
#pragma unroll
	for (int q = 0; q < 7; q++) {
		int xa = max(i0 - sw7_mx0(q), 0), xb = min(i1 + sw7_mx1(q), nx);
		int ya = max(j0 - sw7_my0(q), 0), yb = min(j1 + sw7_my1(q), ny);
		int rows = xb - xa;
		int i, j;
		bool ok;
		if (q == 6) {
			int wdt = yb - ya;
			int ii = r / wdt;
			i = xa + ii;
			j = ya + (r - ii * wdt);
			ok = ii < rows;
		} else {
			int hw = (yb - ya + 1) >> 1;
			int ii = r / hw;
			i = xa + ii;
			int par = (sw7_lag(q) + pidx) & 1;
			j = ya + ((ya + i + par + (q & 1)) & 1) + 2 * (r - ii * hw);
			ok = ii < rows && j < yb;
		}
		code[q] = ok ? ((i - rx0) << 8 | (j - ry0)) : -1;
	}

	/* region node of this thread for the plane loads */
	const int lrx = t / SW7_RY, lry = t - lrx * SW7_RY;
	const int gi = rx0 + lrx, gj = ry0 + lry;
	const bool lin = t < SW7_PL && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
	const long gn = (long)gi * ny + gj;
	const float *src[8] = {xin, xin + N, xin + 2 * N, xin + 3 * N, L.eC, L.eXY, L.eXZ, L.eYZ};

	/* planes of the first step */
#pragma unroll
	for (int f = 0; f < 8; f++)
		for (int pl = 0; pl <= 1 + sw7_up(f) && pl < nz; pl++)
			if (lin) s[sw7_slot(f, pl) + t] = src[f][pl * sz + gn];

	/* right-hand sides and flags of a step's nodes (the flags are kept as
	 * loaded: nothing may wait for the next step's loads before the step
	 * ends) */
	float rhs[7], nrhs[7];
	unsigned flg[7], nflg[7];
	auto fetch = [&](int kk, float *rr, unsigned *ff) {
#pragma unroll
		for (int q = 0; q < 7; q++) {
			int pl = kk - sw7_lag(q) + pidx;
			int cd = code[q];
			asm volatile("" : "+v"(cd));
			rr[q] = 0.f;
			ff[q] = 0;
			if (cd >= 0 && pl >= 0 && pl < nz) {
				long n = pl * sz + (long)(rx0 + (cd >> 8)) * ny + (ry0 + (cd & 255));
				ff[q] = L.flags[n];
				if (b) rr[q] = b[sw7_comp(q) * N + n];
			}
		}
	};
	fetch(0, rhs, flg);
	__syncthreads();

	for (int kk = 0; kk <= nz + 2; kk += 2) {
		fetch(kk + 2, nrhs, nflg);
		float nv[8][2];
#pragma unroll
		for (int q = 0; q < 7; q++) {
			if (q == SW7_LOAD_PHASE) {
				/* the next step's planes, loaded while the last phases run */
#pragma unroll
				for (int f = 0; f < 8; f++)
#pragma unroll
					for (int e = 0; e < 2; e++) {
						int pl = kk + 2 + sw7_up(f) + e;
						nv[f][e] = (lin && pl < nz) ? src[f][pl * sz + gn] : 0.f;
					}
			}
			const int C = sw7_comp(q);
			int pl = __builtin_amdgcn_readfirstlane(kk - sw7_lag(q) + pidx);
			int cd = code[q];
			/* recompute what derives from the node code in every step
			 * instead of holding it in registers across the march */
			asm volatile("" : "+v"(cd));
			if (cd >= 0 && pl >= 0 && pl < nz) {
				int li = cd >> 8, lj = cd & 255;
				Sw7Acc a{s, li * SW7_RY + lj, li, lj, pl, gxh, gxc, gyh, gyc, gzh, gzc};
				unsigned fl = flg[q];
				bool upd = (fl & FL_EX(C)) && !(fl & FL_BC(C));
				if (q < 6) {
					if (upd) {
						float v;
						if (C == 1) v = vel_relax_a<float, 1>(a, rhs[q], v0koef);
						else if (C == 2) v = vel_relax_a<float, 2>(a, rhs[q], v0koef);
						else v = vel_relax_a<float, 3>(a, rhs[q], v0koef);
						s[sw7_slot(C, pl) + a.o] = v;
					}
				} else {
					float pv = a.at(0, 0, 0, 0);
					if (upd) pv = p_relax_a<float>(a, pv, rhs[q], p0koef);
					long n = pl * sz + (long)(rx0 + li) * ny + (ry0 + lj);
					xout[n] = pv;
					xout[N + n] = a.at(1, 0, 0, 0);
					xout[2 * N + n] = a.at(2, 0, 0, 0);
					xout[3 * N + n] = a.at(3, 0, 0, 0);
				}
			}
			sw7_sync();
		}

#pragma unroll
		for (int f = 0; f < 8; f++)
#pragma unroll
			for (int e = 0; e < 2; e++) {
				int pl = kk + 2 + sw7_up(f) + e;
				if (lin && pl < nz) s[sw7_slot(f, pl) + t] = nv[f][e];
			}
#pragma unroll
		for (int q = 0; q < 7; q++) {
			rhs[q] = nrhs[q];
			flg[q] = nflg[q];
		}
		sw7_sync();
	}
}

/* ---- one launch per component: both colours through LDS ------------------ */

/* One FP32 sweep in three launches: L = 1 relaxes vx (colour 0, then 1),
 * L = 2 vy, L = 3 vz and then P; equal to the seven phase launches to the
 * bit.  Launch L reads the sweep's input xa for its own component, P and
 * the later components, and xb (where the earlier launches wrote) for the
 * earlier components; it writes its component(s) to xb.
 *
 * A work-group owns an (x, y) tile of at most SW3_TX x SW3_TY nodes and a
 * z-slab, and marches up one plane per step: colour 0 at plane kk, colour 1
 * (and P for L = 3) at kk-1.  A second-colour row reads the first colour one
 * plane up and down, so one plane of lag is enough; P reads vz at its plane
 * and the one below.  Colour 0 runs on the tile plus the ring that colour 1
 * reads (one line on every side; for L = 3 two below and one above, as P
 * reads vz at (i-1, j-1)), and on one plane below and above the slab (two
 * below for L = 3); those values are the neighbouring work-groups' own,
 * computed by the same function from the same inputs.  The operands of the
 * staged region (tile + 2 lines around, 3 below for L = 3) live in LDS in
 * one ring of planes per field read by the launch; the next plane, the
 * right-hand sides and the flags of the next step are loaded into
 * registers while the current step computes, and the barriers wait for LDS
 * only. */
#define SW3_TX 23
#define SW3_TY 23
#define SW3_RX (SW3_TX + 5)
#define SW3_RY (SW3_TY + 5)
#define SW3_PL (SW3_RX * SW3_RY)
#define SW3_THREADS 512
#define SW3_IT 2 /* nodes a thread relaxes per phase at most */

/* ring depth of field f (0 P, 1 vx, 2 vy, 3 vz, 4 eC, 5 eXY, 6 eXZ, 7 eYZ;
 * 0: not read) in launch L: at least the planes read below the step's plane
 * kk, the planes read above (sw3_up) and kk itself, rounded up to a power of
 * two so that a plane's slot is a mask, not a division */
__host__ __device__ constexpr int sw3_depth(int L, int f)
{
	return L == 1 ? (f == 0 ? 2 : f == 1 ? 4 : f == 2 ? 2 : f == 3 ? 4 : f == 4 ? 2 : f == 5 ? 2 : f == 6 ? 4 : 0)
	     : L == 2 ? (f == 0 ? 2 : f == 1 ? 2 : f == 2 ? 4 : f == 3 ? 4 : f == 4 ? 2 : f == 5 ? 2 : f == 6 ? 0 : 4)
		      : (f == 0 ? 4 : f == 1 ? 4 : f == 2 ? 4 : f == 3 ? 4 : f == 4 ? 4 : f == 5 ? 0 : f == 6 ? 2 : 2);
}
/* planes read below kk: depth - up - 1 before the rounding */
__host__ __device__ constexpr int sw3_dn(int L, int f)
{
	return L == 1 ? (f == 1 ? 2 : f == 2 || f == 3 || f == 5 || f == 6 ? 1 : 0)
	     : L == 2 ? (f == 2 ? 2 : f == 1 || f == 3 || f == 5 || f == 7 ? 1 : 0)
		      : (f == 1 || f == 2 || f == 3 ? 2 : f == 0 || f == 4 || f == 6 || f == 7 ? 1 : 0);
}
__host__ __device__ constexpr int sw3_up(int L, int f)
{
	return L == 1 ? (f == 0 || f == 1 || f == 3 || f == 4 || f == 6)
	     : L == 2 ? (f == 0 || f == 2 || f == 3 || f == 4 || f == 7)
		      : (f == 0 || f == 3 || f == 4);
}
__host__ __device__ constexpr int sw3_base(int L, int f)
{
	return (f > 0 ? sw3_depth(L, 0) : 0) + (f > 1 ? sw3_depth(L, 1) : 0) + (f > 2 ? sw3_depth(L, 2) : 0) +
	       (f > 3 ? sw3_depth(L, 3) : 0) + (f > 4 ? sw3_depth(L, 4) : 0) + (f > 5 ? sw3_depth(L, 5) : 0) +
	       (f > 6 ? sw3_depth(L, 6) : 0) + (f > 7 ? sw3_depth(L, 7) : 0);
}
__host__ __device__ constexpr int sw3_slots(int L) { return sw3_base(L, 8); }
__device__ __forceinline__ int sw3_slot(int L, int f, int pl)
{
	return (sw3_base(L, f) + (int)((unsigned)pl & (unsigned)(sw3_depth(L, f) - 1))) * SW3_PL;
}
/* region margins of phase q (x below, x above, y below, y above), lags and
 * relaxed component; the staged region starts sw3_h(L) lines below */
__host__ __device__ constexpr int sw3_h(int L) { return L == 3 ? 3 : 2; }
__host__ __device__ constexpr int sw3_nq(int L) { return L == 3 ? 3 : 2; }
__host__ __device__ constexpr int sw3_lag(int q) { return q == 0 ? 0 : 1; }
__host__ __device__ constexpr int sw3_m0(int L, int q) { return q == 0 ? (L == 3 ? 2 : 1) : q == 1 ? (L == 3 ? 1 : 0) : 0; }
__host__ __device__ constexpr int sw3_m1(int L, int q) { return q == 0 ? 1 : 0; }

struct Sw3Acc {
	const float *s;
	int L, o, li, lj, pl;
	const float *gxh, *gxc, *gyh, *gyc, *gzh, *gzc;
	/* o is the node's offset less one line and one row, so that every
	 * stencil offset below is a non-negative constant (an LDS immediate) */
	__device__ __forceinline__ float at(int f, int di, int dj, int dk) const
	{
		const float *base = s + sw3_slot(L, f, pl + dk) + o;
		return base[(di + 1) * SW3_RY + (dj + 1)];
	}
	template <int C> __device__ __forceinline__ float u(int di, int dj, int dk) const { return at(C, di, dj, dk); }
	__device__ __forceinline__ float eC(int di, int dj, int dk) const { return at(4, di, dj, dk); }
	__device__ __forceinline__ float eXY(int di, int dj, int dk) const { return at(5, di, dj, dk); }
	__device__ __forceinline__ float eXZ(int di, int dj, int dk) const { return at(6, di, dj, dk); }
	__device__ __forceinline__ float eYZ(int di, int dj, int dk) const { return at(7, di, dj, dk); }
	__device__ __forceinline__ float ihx(int d) const { return gxh[li + d]; }
	__device__ __forceinline__ float icx(int d) const { return gxc[li + d]; }
	__device__ __forceinline__ float ihy(int d) const { return gyh[lj + d]; }
	__device__ __forceinline__ float icy(int d) const { return gyc[lj + d]; }
	__device__ __forceinline__ float ihz(int d) const { return gzh[pl + d]; }
	__device__ __forceinline__ float icz(int d) const { return gzc[pl + d]; }
};

struct Sw3Geom {
	int ntx, nty, nzs; /* tiles in x, y; slabs in z */
	int remap;         /* XCD-contiguous work order */
};

static inline size_t sw3_lds_bytes(int L, int nz)
{
	return sizeof(float) * ((size_t)sw3_slots(L) * SW3_PL + 2 * SW3_RX + 2 * SW3_RY + 2 * (size_t)nz);
}

template <int L>
__global__ void __launch_bounds__(SW3_THREADS, 4)
k_sweep3(LevelView<float> V, const float *__restrict__ xa, const float *__restrict__ xb_in, float *__restrict__ xout,
	 const float *__restrict__ b, Sw3Geom G, float v0koef, float p0koef)
{
	extern __shared__ __align__(16) unsigned char lds_raw[];
	float *s = (float *)lds_raw;
	float *gxh = s + sw3_slots(L) * SW3_PL, *gxc = gxh + SW3_RX, *gyh = gxc + SW3_RX, *gyc = gyh + SW3_RY;
	float *gzh = gyc + SW3_RY, *gzc = gzh + V.nz;
	constexpr int NQ = sw3_nq(L);
	int w, dummy;
	if (!block_work(G.remap, G.ntx * G.nty * G.nzs, 1, 0, &w, &dummy)) return;
	const int tx = w % G.ntx, rest = w / G.ntx;
	const int ty = rest % G.nty, sl = rest / G.nty;
	const int nx = V.nx, ny = V.ny, nz = V.nz;
	const long N = V.N, sz = (long)nx * ny;
	const int i0 = (int)((long)tx * nx / G.ntx), i1 = (int)((long)(tx + 1) * nx / G.ntx);
	const int j0 = (int)((long)ty * ny / G.nty), j1 = (int)((long)(ty + 1) * ny / G.nty);
	const int k0 = (int)((long)sl * nz / G.nzs), k1 = (int)((long)(sl + 1) * nz / G.nzs);
	const int rx0 = i0 - sw3_h(L), ry0 = j0 - sw3_h(L);
	const int t = threadIdx.x;

	for (int q = t; q < SW3_RX; q += SW3_THREADS) {
		int i = rx0 + q;
		bool in = i >= 0 && i < nx;
		gxh[q] = in ? V.ihx[i] : 0.f;
		gxc[q] = in ? V.icx[i] : 0.f;
	}
	for (int q = t; q < SW3_RY; q += SW3_THREADS) {
		int j = ry0 + q;
		bool in = j >= 0 && j < ny;
		gyh[q] = in ? V.ihy[j] : 0.f;
		gyc[q] = in ? V.icy[j] : 0.f;
	}
	for (int q = t; q < nz; q += SW3_THREADS) {
		gzh[q] = V.ihz[q];
		gzc[q] = V.icz[q];
	}

	/* this thread's nodes of each phase: region offsets li << 8 | lj (or
	 * 0xffff: none) for even planes in the low and odd planes in the high
	 * 16 bits */
	unsigned code[NQ][SW3_IT];
#pragma unroll
	for (int q = 0; q < NQ; q++) {
		int xa0 = max(i0 - sw3_m0(L, q), 0), xb0 = min(i1 + sw3_m1(L, q), nx);
		int ya0 = max(j0 - sw3_m0(L, q), 0), yb0 = min(j1 + sw3_m1(L, q), ny);
		int rows = xb0 - xa0;
#pragma unroll
		for (int it = 0; it < SW3_IT; it++) code[q][it] = 0;
#pragma unroll
		for (int par = 0; par < 2; par++)
#pragma unroll
			for (int it = 0; it < SW3_IT; it++) {
				int r = t + it * SW3_THREADS;
				int i, j;
				bool ok;
				if (q == 2) {
					int wdt = yb0 - ya0, ii = r / wdt;
					i = xa0 + ii;
					j = ya0 + (r - ii * wdt);
					ok = ii < rows;
				} else {
					int hw = (yb0 - ya0 + 1) >> 1, ii = r / hw;
					i = xa0 + ii;
					j = ya0 + ((ya0 + i + par + q) & 1) + 2 * (r - ii * hw);
					ok = ii < rows && j < yb0;
				}
				code[q][it] |= (ok ? (unsigned)((i - rx0) << 8 | (j - ry0)) : 0xffffu) << (16 * par);
			}
	}

	/* region node of this thread for the plane loads (up to two) */
	auto src = [&](int f) -> const float * {
		switch (f) {
		case 0: return xa;
		case 1: return L >= 2 ? xb_in + N : xa + N;
		case 2: return L >= 3 ? xb_in + 2 * N : xa + 2 * N;
		case 3: return xa + 3 * N;
		case 4: return V.eC;
		case 5: return V.eXY;
		case 6: return V.eXZ;
		default: return V.eYZ;
		}
	};
	long gn[2];
	bool lin[2];
#pragma unroll
	for (int e = 0; e < 2; e++) {
		int r = t + e * SW3_THREADS;
		int lrx = r / SW3_RY, lry = r - lrx * SW3_RY;
		int gi = rx0 + lrx, gj = ry0 + lry;
		lin[e] = r < SW3_PL && gi >= 0 && gi < nx && gj >= 0 && gj < ny;
		gn[e] = (long)gi * ny + gj;
	}

	const int kb = k0 - (L == 3 ? 2 : 1); /* first step: colour 0 on the plane(s) below the slab */
	const int ke = k1;                     /* last step: colour 1 / P on the slab's top plane */
	/* planes of the first step */
#pragma unroll
	for (int f = 0; f < 8; f++) {
		if (!sw3_depth(L, f)) continue;
		for (int pl = max(kb - sw3_dn(L, f), 0); pl <= kb + sw3_up(L, f) && pl < nz; pl++)
#pragma unroll
			for (int e = 0; e < 2; e++)
				if (lin[e]) s[sw3_slot(L, f, pl) + t + e * SW3_THREADS] = src(f)[pl * sz + gn[e]];
	}

	/* right-hand sides and flags of a step's nodes, loaded a step ahead */
	float rhs[NQ][SW3_IT], nrhs[NQ][SW3_IT];
	unsigned flg[NQ][SW3_IT], nflg[NQ][SW3_IT];
	auto fetch = [&](int kk, float (*rr)[SW3_IT], unsigned (*ff)[SW3_IT]) {
#pragma unroll
		for (int q = 0; q < NQ; q++) {
			int pl = kk - sw3_lag(q);
			int par = pl & 1;
#pragma unroll
			for (int it = 0; it < SW3_IT; it++) {
				unsigned cd = (code[q][it] >> (16 * par)) & 0xffffu;
				asm volatile("" : "+v"(cd));
				rr[q][it] = 0.f;
				ff[q][it] = 0;
				if (cd != 0xffffu && pl >= 0 && pl < nz) {
					long n = pl * sz + (long)(rx0 + (cd >> 8)) * ny + (ry0 + (cd & 255));
					ff[q][it] = V.flags[n];
					if (b) rr[q][it] = b[(q == 2 ? 0 : L) * N + n];
				}
			}
		}
	};
	fetch(kb, rhs, flg);
	__syncthreads();

	for (int kk = kb; kk <= ke; kk++) {
		fetch(kk + 1, nrhs, nflg);
		float nv[8][2];
#pragma unroll
		for (int f = 0; f < 8; f++)
#pragma unroll
			for (int e = 0; e < 2; e++) {
				int pl = kk + 1 + sw3_up(L, f);
				nv[f][e] = (sw3_depth(L, f) && lin[e] && pl < nz) ? src(f)[pl * sz + gn[e]] : 0.f;
			}
#pragma unroll
		for (int q = 0; q < NQ; q++) {
			const int C = q == 2 ? 0 : L;
			const int pl = kk - sw3_lag(q);
			/* colour 0 on the slab's planes and the halo planes around
			 * it; colour 1 and P on the slab's own planes, and for L = 3
			 * colour 1 also on the plane below (P reads vz there) */
			const int lo = q == 0 ? kb : (q == 1 && L == 3 ? k0 - 1 : k0);
			const int hi = q == 0 ? ke : k1 - 1;
			if (pl >= lo && pl <= hi && pl >= 0 && pl < nz) {
				const int par = pl & 1;
#pragma unroll
				for (int it = 0; it < SW3_IT; it++) {
					unsigned cd = (code[q][it] >> (16 * par)) & 0xffffu;
					asm volatile("" : "+v"(cd));
					if (cd == 0xffffu) continue;
					int li = cd >> 8, lj = cd & 255;
					Sw3Acc a{s, L, (li - 1) * SW3_RY + lj - 1, li, lj, pl, gxh, gxc, gyh, gyc, gzh, gzc};
					unsigned fl = flg[q][it];
					bool upd = (fl & FL_EX(C)) && !(fl & FL_BC(C));
					if (q == 2) {
						/* P: every node of the tile; the new value goes
						 * straight out (no later phase reads it) */
						float v = a.at(0, 0, 0, 0);
						if (upd) v = p_relax_a<float>(a, v, rhs[q][it], p0koef);
						xout[pl * sz + (long)(rx0 + li) * ny + (ry0 + lj)] = v;
					} else if (upd) {
						float v;
						if (L == 1) v = vel_relax_a<float, 1>(a, rhs[q][it], v0koef);
						else if (L == 2) v = vel_relax_a<float, 2>(a, rhs[q][it], v0koef);
						else v = vel_relax_a<float, 3>(a, rhs[q][it], v0koef);
						s[sw3_slot(L, C, pl) + li * SW3_RY + lj] = v;
					}
				}
			}
			sw7_sync();
		}
		/* the tile's component at plane kk-1 is final: out (its ring slot is
		 * not the one the loads below overwrite) */
		if (kk - 1 >= k0 && kk - 1 < k1) {
			const int pl = kk - 1, wdt = j1 - j0, cnt = (i1 - i0) * wdt;
			for (int r = t; r < cnt; r += SW3_THREADS) {
				int ii = r / wdt, i = i0 + ii, j = j0 + (r - ii * wdt);
				xout[L * N + pl * sz + (long)i * ny + j] =
					s[sw3_slot(L, L, pl) + (i - rx0) * SW3_RY + (j - ry0)];
			}
		}
#pragma unroll
		for (int f = 0; f < 8; f++)
#pragma unroll
			for (int e = 0; e < 2; e++) {
				int pl = kk + 1 + sw3_up(L, f);
				if (sw3_depth(L, f) && lin[e] && pl < nz) s[sw3_slot(L, f, pl) + t + e * SW3_THREADS] = nv[f][e];
			}
#pragma unroll
		for (int q = 0; q < NQ; q++)
#pragma unroll
			for (int it = 0; it < SW3_IT; it++) {
				rhs[q][it] = nrhs[q][it];
				flg[q][it] = nflg[q][it];
				/* wait for them here, where they have long arrived: their
				 * loads are conditional, so a wait placed at their use in
				 * the next step would also drain that step's prefetch */
				asm volatile("" : "+v"(rhs[q][it]), "+v"(flg[q][it]));
			}
		sw7_sync();
	}
}

/* Boundary rows x_u = c + k x_partner for the records [start, end) of one
 * dependency depth (partners of depth-d rows are interior or depth < d). */
template <typename T>
__device__ __forceinline__ void bc_update(T *__restrict__ x, const BcRec<T> &bc, int t, int homog)
{
	int p = bc.partner[t];
	T v = homog ? (T)0 : bc.c[t];
	if (p >= 0) v += bc.k[t] * x[p];
	x[bc.u[t]] = v;
}

template <typename T>
__global__ void k_bc(T *__restrict__ x, BcRec<T> bc, int start, int end, int homog)
{
	int t = start + blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= end) return;
	bc_update<T>(x, bc, t, homog);
}

/* All BC records [start, end) in one launch: a record whose partner is
 * itself a BC row follows the partner chain (at most BC_MAXCHAIN records,
 * checked on the host) down to an interior unknown and evaluates it
 * innermost first, each step the expression of bc_update, so every value
 * equals what the depth-by-depth launches store.  Only interior unknowns
 * are read, only BC unknowns written: no ordering between threads. */
#define BC_MAXCHAIN 8
template <typename T>
__global__ void k_bc_chain(T *__restrict__ x, BcRec<T> bc, const int *__restrict__ bcid, int start, int end,
			   int homog)
{
	int t = start + blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= end) return;
	int chain[BC_MAXCHAIN];
	int m = 0, r = t;
	for (;;) {
		chain[m++] = r;
		int p = bc.partner[r];
		if (p < 0 || m == BC_MAXCHAIN) break;
		int rp = bcid[p];
		if (rp < 0) break;
		r = rp;
	}
	int ri = chain[m - 1], pi = bc.partner[ri];
	T v = homog ? (T)0 : bc.c[ri];
	if (pi >= 0) v += bc.k[ri] * x[pi];
	for (int q = m - 2; q >= 0; q--) {
		int rq = chain[q];
		T w = homog ? (T)0 : bc.c[rq];
		w += bc.k[rq] * v;
		v = w;
	}
	x[bc.u[t]] = v;
}

/* Per-axis transfer maps between a fine and a coarse level for the three
 * staggering kinds (0 node line, 1 mid-point, 2 cell centre), built on the
 * host from the m10min/m20min/m30min rules of vprestrict/vpprolong
 * (move3mg.c:2776-2789, 3201-3214). Forward (per fine index): base, wlo,
 * whi.  Inverse (per coarse index): CSR list of (fine index, weight) in
 * ascending fine index. */
template <typename T> struct AxisMap {
	const int *base[3];
	const T *wlo[3], *whi[3];
	const int *istart[3];
	const int *iidx[3];
	const T *iw[3];
};

template <typename T> struct XferView {
	AxisMap<T> ax[3]; /* x, y, z */
};

__device__ __constant__ int c_kind[4][3] = {{2, 2, 2}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

/* Longest inverse list of one axis (fine lines gathered by one coarse
 * line); the host checks the maps against it. */
#define XFER_MAXL 6

/* Residual restriction in gather form, one thread per coarse unknown
 * (component c = blockIdx.z).  The coarse line's y-list is held in
 * registers and the fine flag/value loads of a whole y-list are issued
 * together.  Accumulation order equals the reference's serial scatter
 * (fine nodes in ascending index) and nothing is contracted, so in FP64
 * the result is the reference's to the bit.  Coarse x is zeroed
 * (move3mg.c:3508-3518). */
template <typename T>
__device__ __forceinline__ void restrict_node(const LevelView<T> &F, const LevelView<T> &Cl, const XferView<T> &M,
					      const T *__restrict__ rf, T *__restrict__ bc_out, T *__restrict__ xc,
					      int I, int J, int K, int c)
{
#pragma clang fp contract(off)
	long Nc = Cl.N, Nf = F.N;
	long nc = (long)K * Cl.nx * Cl.ny + (long)I * Cl.ny + J;
	unsigned flc = Cl.flags[nc];
	T acc = 0;
	if ((flc & FL_EX(c)) && !(flc & FL_BC(c))) {
		int kx = c_kind[c][0], ky = c_kind[c][1], kz = c_kind[c][2];
		int z0 = M.ax[2].istart[kz][K], z1 = M.ax[2].istart[kz][K + 1];
		int x0 = M.ax[0].istart[kx][I], x1 = M.ax[0].istart[kx][I + 1];
		int y0 = M.ax[1].istart[ky][J], ny = M.ax[1].istart[ky][J + 1] - y0;
		long sxf = F.ny, szf = (long)F.nx * F.ny;
		int jf[XFER_MAXL];
		T wy[XFER_MAXL];
#pragma unroll
		for (int e = 0; e < XFER_MAXL; e++) {
			jf[e] = e < ny ? M.ax[1].iidx[ky][y0 + e] : 0;
			wy[e] = e < ny ? M.ax[1].iw[ky][y0 + e] : (T)0;
		}
		const T *r = rf + c * Nf;
		T ndc = (T)(float)Cl.eC[nc];
		T sw = 0;
		for (int a = z0; a < z1; a++) {
			long kf = M.ax[2].iidx[kz][a];
			T wz = M.ax[2].iw[kz][a];
			for (int bq = x0; bq < x1; bq++) {
				long row = kf * szf + (long)M.ax[0].iidx[kx][bq] * sxf;
				T wx = M.ax[0].iw[kx][bq];
				unsigned fl[XFER_MAXL];
				T q[XFER_MAXL], ef[XFER_MAXL];
#pragma unroll
				for (int e = 0; e < XFER_MAXL; e++) {
					fl[e] = 0;
					q[e] = 0;
					ef[e] = 0;
					if (e < ny) {
						fl[e] = F.flags[row + jf[e]];
						q[e] = r[row + jf[e]];
						if (c == 0) ef[e] = F.eC[row + jf[e]];
					}
				}
#pragma unroll
				for (int e = 0; e < XFER_MAXL; e++) {
					if (!(fl[e] & FL_EX(c)) || (fl[e] & FL_BC(c))) continue;
					T swt = wx * wy[e] * wz;
					if (c == 0)
						acc += q[e] * (T)(float)ef[e] / ndc * swt;
					else
						acc += q[e] * swt;
					sw += swt;
				}
			}
		}
		if (sw != (T)0) acc /= sw;
	}
	bc_out[c * Nc + nc] = acc;
	xc[c * Nc + nc] = 0;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_restrict(LevelView<T> F, LevelView<T> Cl, XferView<T> M, const T *__restrict__ rf,
	   T *__restrict__ bc_out, T *__restrict__ xc)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int I = t / Cl.ny, J = t - I * Cl.ny;
	if (I >= Cl.nx) return;
	restrict_node<T>(F, Cl, M, rf, bc_out, xc, I, J, blockIdx.y, blockIdx.z);
}

/* Correction prolongation x_f += P x_c (move3mg.c:2760-3060), per fine
 * node, the four unknown kinds in one thread (one flag load); the eight
 * coarse values of a kind are loaded before they are summed.  v1koef for
 * velocities, p1koef for pressure; the reference's term order, nothing
 * contracted. */
template <typename T>
__device__ __forceinline__ void prolong_node(const LevelView<T> &F, const LevelView<T> &Cl, const XferView<T> &M,
					     const T *__restrict__ xc, T *__restrict__ xf, int i, int j, int k,
					     T v1koef, T p1koef)
{
#pragma clang fp contract(off)
	long Nc = Cl.N, Nf = F.N;
	long nf = (long)k * F.nx * F.ny + (long)i * F.ny + j;
	unsigned fl = F.flags[nf];
	long sxc = Cl.ny, szc = (long)Cl.nx * Cl.ny;
	for (int c = 0; c < 4; c++) {
		if (!(fl & FL_EX(c))) continue;
		int kx = c_kind[c][0], ky = c_kind[c][1], kz = c_kind[c][2];
		int bx = M.ax[0].base[kx][i], by = M.ax[1].base[ky][j], bz = M.ax[2].base[kz][k];
		T wx[2] = {M.ax[0].wlo[kx][i], M.ax[0].whi[kx][i]};
		T wy[2] = {M.ax[1].wlo[ky][j], M.ax[1].whi[ky][j]};
		T wz[2] = {M.ax[2].wlo[kz][k], M.ax[2].whi[kz][k]};
		T kf = c ? v1koef : p1koef;
		const T *xcc = xc + c * Nc;
		T v[8];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int bq = 0; bq < 2; bq++)
#pragma unroll
				for (int e = 0; e < 2; e++)
					v[a * 4 + bq * 2 + e] = xcc[(bz + a) * szc + (bx + bq) * sxc + (by + e)];
		T acc = xf[c * Nf + nf];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int bq = 0; bq < 2; bq++)
#pragma unroll
				for (int e = 0; e < 2; e++) {
					T swt = wx[bq] * wy[e] * wz[a];
					acc += v[a * 4 + bq * 2 + e] * swt * kf;
				}
		xf[c * Nf + nf] = acc;
	}
}

template <typename T>
__global__ void __launch_bounds__(256)
k_prolong(LevelView<T> F, LevelView<T> Cl, XferView<T> M, const T *__restrict__ xc,
	  T *__restrict__ xf, T v1koef, T p1koef)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int i = t / F.ny, j = t - i * F.ny;
	if (i >= F.nx) return;
	prolong_node<T>(F, Cl, M, xc, xf, i, j, blockIdx.y, v1koef, p1koef);
}

/* ---- the coarsest level in one work-group, from LDS --------------------- */

/* The coarsest level (12^3 in the usual hierarchies) is smoothed nu_coarse
 * times per V-cycle: as launches, each colour pass, pressure pass and BC
 * depth costs a few microseconds for well under a microsecond of work.
 * k_coarsest runs all of its sweeps in one work-group with the unknowns,
 * the four viscosity planes and the flags held in LDS (the right-hand side,
 * line factors and BC records stay in global memory): every launch becomes
 * a work-group-strided loop over the same per-node functions followed by a
 * barrier, in launch order, so the result is the launched sweeps' bit for
 * bit. */
#define COARSEST_THREADS 1024
#define COARSEST_MAXD 8

/* ---- coupled cell smoother (SCGS, opt-in) --------------------------------
 *
 * Symmetric coupled Gauss-Seidel (Vanka's box smoother with the velocity
 * block reduced to its diagonal): cell P(i,j,k) and its six faces
 * vx(i,j-1,k-1) (+x), vx(i-1,j-1,k-1) (-x), vy(i-1,j,k-1) (+y),
 * vy(i-1,j-1,k-1) (-y), vz(i-1,j-1,k) (+z), vz(i-1,j-1,k-1) (-z) are
 * corrected together: with r_f, A_f the face rows' residuals and diagonals,
 * G_f the rows' coefficient of this cell's P and D_f the continuity row's
 * coefficient of the face,
 *   dp = (sum_f D_f r_f / A_f - r_c) / sum_f D_f G_f / A_f,
 *   du_f = (r_f - G_f dp) / A_f.
 * BC and missing unknowns take no part.  Cells in two colours (i+j+k
 * parity; two cells of one colour share no face): k_scgs_cell computes the
 * seven corrections of every cell of the colour from the current iterate
 * into D (7 planes: dp, +x, -x, +y, -y, +z, -z), k_scgs_apply adds them,
 * damped, each face from its one cell of that colour. */
template <typename T>
__global__ void __launch_bounds__(256)
k_scgs_cell(LevelView<T> L, const T *__restrict__ x, const T *__restrict__ b, T *__restrict__ D, int colour,
	    int remap, int chunks)
{
	/* one thread per cell of the colour: (i, k) and every other j; blocks
	 * over (chunk, plane) as block_work maps them */
	int ch, k;
	if (!block_work(remap, chunks, L.nz, 0, &ch, &k)) return;
	int t = ch * blockDim.x + threadIdx.x;
	int hy = (L.ny + 1) >> 1;
	int i = t / hy, tj = t - i * hy;
	if (i >= L.nx) return;
	int j = 2 * tj + ((i + k + colour) & 1);
	if (j >= L.ny) return;
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	long n = (long)k * sz + (long)i * sx + j;
	T dp = 0, du[6] = {0, 0, 0, 0, 0, 0};
	unsigned fc = L.flags[n];
	if ((fc & FL_EX(0)) && !(fc & FL_BC(0)) && i >= 1 && j >= 1 && k >= 1) {
		const T *P = x, *vx = x + N, *vy = x + 2 * N, *vz = x + 3 * N;
		T rc = (b ? b[n] : (T)0) - row_ax<T, T>(L, P, vx, vy, vz, 0, i, j, k, (T *)0);
		const T Dc[6] = {L.ihx[i - 1], -L.ihx[i - 1], L.ihy[j - 1], -L.ihy[j - 1], L.ihz[k - 1], -L.ihz[k - 1]};
		const T Gc[6] = {(T)2 * L.icx[i], -(T)2 * L.icx[i - 1], (T)2 * L.icy[j], -(T)2 * L.icy[j - 1],
				 (T)2 * L.icz[k], -(T)2 * L.icz[k - 1]};
		T rf[6], af[6], S = 0, R = 0;
#pragma unroll
		for (int f = 0; f < 6; f++) {
			const int comp = 1 + f / 2;
			const int fi = i - (f == 0 ? 0 : 1), fj = j - (f == 2 ? 0 : 1), fk = k - (f == 4 ? 0 : 1);
			long nf = (long)fk * sz + (long)fi * sx + fj;
			unsigned ff = L.flags[nf];
			rf[f] = 0;
			af[f] = 0;
			if ((ff & FL_EX(comp)) && !(ff & FL_BC(comp))) {
				T d;
				T ax = row_ax<T, T>(L, P, vx, vy, vz, comp, fi, fj, fk, &d);
				rf[f] = (b ? b[comp * N + nf] : (T)0) - ax;
				af[f] = d;
				S += Dc[f] * Gc[f] / d;
				R += Dc[f] * rf[f] / d;
			}
		}
		if (S != 0) dp = (R - rc) / S;
#pragma unroll
		for (int f = 0; f < 6; f++)
			if (af[f] != 0) du[f] = (rf[f] - Gc[f] * dp) / af[f];
	}
	D[n] = dp;
#pragma unroll
	for (int f = 0; f < 6; f++) D[(1 + f) * N + n] = du[f];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_scgs_apply(LevelView<T> L, T *__restrict__ x, const T *__restrict__ D, int colour, T wv, T wp, int remap,
	     int chunks)
{
	int ch, k;
	if (!block_work(remap, chunks, L.nz, 0, &ch, &k)) return;
	int t = ch * blockDim.x + threadIdx.x;
	int i = t / L.ny, j = t - i * L.ny;
	if (i >= L.nx) return;
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	long n = (long)k * sz + (long)i * sx + j;
	unsigned fl = L.flags[n];
	const bool same = ((i + j + k) & 1) == colour;
	if (same && (fl & FL_EX(0)) && !(fl & FL_BC(0))) x[n] += wp * D[n];
#pragma unroll
	for (int comp = 1; comp <= 3; comp++) {
		if (!(fl & FL_EX(comp)) || (fl & FL_BC(comp))) continue;
		/* the face's cell of this colour: the one on its + side (this
		 * face is its - face) or on its - side (its + face) */
		int ci = i + 1, cj = j + 1, ck = k + 1, slot = 2 * comp;
		if (same) {
			if (comp == 1) ci = i;
			else if (comp == 2) cj = j;
			else ck = k;
			slot = 2 * comp - 1;
		}
		if (ci >= L.nx || cj >= L.ny || ck >= L.nz) continue;
		long nc = (long)ck * sz + (long)ci * sx + cj;
		x[comp * N + n] += wv * D[slot * N + nc];
	}
}

template <typename T> struct CoarsestArgs {
	LevelView<T> v;
	T *x;
	const T *b;
	const T *eta;                    /* eC, eXY, eXZ, eYZ planes */
	BcRec<T> bc;
	int nd;                          /* BC depth ranges 1..nd */
	int ds[COARSEST_MAXD + 2];
	int sweeps;
	T v0koef, p0koef;
};

template <typename T>
__device__ void wg_sweep(const LevelView<T> &L, T *x, const T *b, const CoarsestArgs<T> &A)
{
	int hy = (L.ny + 1) >> 1;
	int plane = L.nx * hy;
	int tot = L.nz * plane;
#pragma unroll
	for (int comp = 1; comp <= 3; comp++)
		for (int col = 0; col < 2; col++) {
			for (int t = threadIdx.x; t < tot; t += COARSEST_THREADS) {
				int k = t / plane, rem = t - k * plane;
				int i = rem / hy, tj = rem - i * hy;
				int j = 2 * tj + ((i + k + col) & 1);
				if (j >= L.ny) continue;
				if (comp == 1) vel_update<T, 1>(L, x, b, i, j, k, A.v0koef);
				else if (comp == 2) vel_update<T, 2>(L, x, b, i, j, k, A.v0koef);
				else vel_update<T, 3>(L, x, b, i, j, k, A.v0koef);
			}
			__syncthreads();
		}
	int pl = L.nx * L.ny, N = (int)L.N;
	for (int t = threadIdx.x; t < N; t += COARSEST_THREADS) {
		int k = t / pl, rem = t - k * pl;
		int i = rem / L.ny, j = rem - i * L.ny;
		p_update<T>(L, x, b, i, j, k, A.p0koef);
	}
	__syncthreads();
	for (int d = 1; d <= A.nd; d++) {
		for (int t = A.ds[d] + threadIdx.x; t < A.ds[d + 1]; t += COARSEST_THREADS) bc_update<T>(x, A.bc, t, 1);
		__syncthreads();
	}
}

/* dynamic LDS: 4N unknowns, 4N viscosities (T), N flag bytes */
template <typename T>
__global__ void __launch_bounds__(COARSEST_THREADS) k_coarsest(CoarsestArgs<T> A)
{
	extern __shared__ __align__(16) unsigned char lds[];
	int N = (int)A.v.N;
	T *sx = (T *)lds, *se = sx + 4 * N;
	uint8_t *sf = (uint8_t *)(se + 4 * N);
	for (int t = threadIdx.x; t < 4 * N; t += COARSEST_THREADS) {
		sx[t] = A.x[t];
		se[t] = A.eta[t];
	}
	for (int t = threadIdx.x; t < N; t += COARSEST_THREADS) sf[t] = A.v.flags[t];
	__syncthreads();
	LevelView<T> L = A.v;
	L.eC = se;
	L.eXY = se + N;
	L.eXZ = se + 2 * N;
	L.eYZ = se + 3 * N;
	L.flags = sf;
	for (int s = 0; s < A.sweeps; s++) wg_sweep<T>(L, sx, A.b, A);
	for (int t = threadIdx.x; t < 4 * N; t += COARSEST_THREADS) A.x[t] = sx[t];
}

/* ---- vector helpers (Krylov, precision conversion) --------------------- */

template <typename T, typename U>
__global__ void k_copy_scale(T *__restrict__ y, const U *__restrict__ x, double a, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) y[t] = (T)((double)x[t] * a);
}

/* y += a x */
template <typename T, typename U>
__global__ void k_axpy(T *__restrict__ y, const U *__restrict__ x, double a, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) y[t] = (T)((double)y[t] + a * (double)x[t]);
}

/* partial dot products: block b sums a fixed chunk */
template <typename T, typename U>
__global__ void __launch_bounds__(256) k_dot_part(const T *__restrict__ a, const U *__restrict__ b,
						  long n, double *partials)
{
	__shared__ double s[256];
	long chunk = (n + gridDim.x - 1) / gridDim.x;
	long lo = (long)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
	double acc = 0;
	for (long t = lo + threadIdx.x; t < hi; t += blockDim.x) acc += (double)a[t] * (double)b[t];
	s[threadIdx.x] = acc;
	__syncthreads();
	for (int q = 128; q > 0; q >>= 1) {
		if (threadIdx.x < q) s[threadIdx.x] += s[threadIdx.x + q];
		__syncthreads();
	}
	if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

/* fixed-order final reduction of `count` groups of `width` partials */
__global__ void k_reduce_final(const double *partials, long nparts, int width, double *out)
{
	__shared__ double s[256];
	for (int w = 0; w < width; w++) {
		double acc = 0;
		for (long t = threadIdx.x; t < nparts; t += blockDim.x) acc += partials[t * width + w];
		s[threadIdx.x] = acc;
		__syncthreads();
		for (int q = 128; q > 0; q >>= 1) {
			if (threadIdx.x < q) s[threadIdx.x] += s[threadIdx.x + q];
			__syncthreads();
		}
		if (threadIdx.x == 0) out[w] = s[0];
		__syncthreads();
	}
}
