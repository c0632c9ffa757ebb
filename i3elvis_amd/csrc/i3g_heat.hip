/*
 * i3g heat path on MI355X: the implicit temperature step of the reference
 * (titerate, heat3mg.c:287-588) kept resident on the device.  Level-0 rows
 * of heaterr/tbonderr are evaluated matrix-free in FP64 and relaxed by
 * red-black Gauss-Seidel (the colour-ordered form of gausmatt's sweep,
 * heat3mg.c:906-961) until the reference's heat norm reaches heatmin.
 * C ABI in include/i3g.h.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/i3g.h"

/* products and sums rounded one at a time, as the reference's x86-64 build */
#pragma clang fp contract(off)

#include "grid_interp.h"

#define TCHK(x)                                                                \
	do {                                                                   \
		hipError_t e_ = (x);                                           \
		if (e_ != hipSuccess) {                                        \
			fprintf(stderr, "i3g heat: HIP error %s at %s:%d\n",  \
				hipGetErrorString(e_), __FILE__, __LINE__);    \
			return I3G_EDEVICE;                                    \
		}                                                              \
	} while (0)

#define NB 256
#define MAXDEPTH 8

struct i3g_heat {
	hipStream_t s = nullptr;
	int nx = 0, ny = 0, nz = 0;
	long N = 0;
	double *gx = nullptr, *gy = nullptr, *gz = nullptr;
	/* boundary rows per node: partner (-2 none, -1 constant only),
	 * constant, coefficient; boundary node lists ordered by dependency
	 * depth (a partner's row is applied before the rows that read it) */
	int *bp = nullptr;
	float *bc = nullptr, *bk = nullptr;
	int *blist = nullptr;
	std::vector<long> boff;   /* boff[d]..boff[d+1]: nodes of depth d */
	/* fields */
	float *kt = nullptr, *rocp = nullptr, *ht = nullptr, *tk0 = nullptr;
	double *T = nullptr, *b = nullptr, *diag = nullptr;
	double *part = nullptr, *res = nullptr;
	std::vector<int> hb;      /* host copy of the T block of bondm */
	/* adiabatic and shear heating of heaterr's right-hand side
	 * (i3g_heat_set_sources): fields on the device and the source as the
	 * reference keeps it in tk1 (float) */
	int adiabyn = 0, frictyn = 0;
	double gk[3] = {0, 0, 0}, ysize = 0, stp100 = 0;
	float *vx = nullptr, *vy = nullptr, *vz = nullptr, *et = nullptr, *ro = nullptr, *se = nullptr;
	double *pr = nullptr;
	float *src = nullptr;
};

template <typename T> static T *halloc(size_t n, int *err)
{
	void *p = nullptr;
	if (n == 0) n = 1;
	if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { *err = I3G_ENOMEM; return nullptr; }
	return (T *)p;
}

/* node index from the (plane, in-plane) launch; y fastest as the reference */
#define NODE_IDX                                                               \
	int t = blockIdx.x * blockDim.x + threadIdx.x;                         \
	int m3 = blockIdx.y;                                                   \
	int m1 = t / ny, m2 = t - m1 * ny;                                     \
	if (m1 >= nx) return;                                                  \
	long sx = ny, sz = (long)nx * ny;                                      \
	long v3 = m3 * sz + m1 * sx + m2;

/* Face conductances of an interior row (heaterr ynerr=1, heat3mg.c:182-220):
 * k_f = (k_a + k_b)/2 with the float sum of the reference, times
 * 2/(g[m+1]-g[m-1]) / (face spacing). */
struct Cond {
	double xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Cond conds(const float *__restrict__ kt, const double *gx, const double *gy,
				      const double *gz, int m1, int m2, int m3, long v3, long sx, long sz)
{
	Cond c;
	float k3 = kt[v3];
	double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
	double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
	double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
	c.xm = (double)(kt[v3 - sx] + k3) / 2.0 * xkf / (gx[m1] - gx[m1 - 1]);
	c.xp = (double)(k3 + kt[v3 + sx]) / 2.0 * xkf / (gx[m1 + 1] - gx[m1]);
	c.ym = (double)(kt[v3 - 1] + k3) / 2.0 * ykf / (gy[m2] - gy[m2 - 1]);
	c.yp = (double)(k3 + kt[v3 + 1]) / 2.0 * ykf / (gy[m2 + 1] - gy[m2]);
	c.zm = (double)(kt[v3 - sz] + k3) / 2.0 * zkf / (gz[m3] - gz[m3 - 1]);
	c.zp = (double)(k3 + kt[v3 + sz]) / 2.0 * zkf / (gz[m3 + 1] - gz[m3]);
	return c;
}

/* heaterr in right-hand-side mode (ynerr=0, heat3mg.c:54-160) at the
 * interior nodes: src = radiogenic heat + adiabatic heating (adiabyn 1:
 * alpha T rho (g.v); 2: alpha T (v.grad P) with the trilinear P gradient
 * between the eight surrounding cell centres) + shear heating (frictyn:
 * sigma:eps, normal terms from the eight centres, shear terms from the two
 * edge nodes along the third axis), the terms only where rho > 2000 and
 * y < ysize - stp100; v at the node is vxyzcalc1's.  Products of two float
 * fields stay float as in the reference.  src is stored as float (tk1);
 * each block writes the largest |src / (rho cp)| of its interior nodes. */
struct HeatSrc {
	int adiabyn, frictyn;
	double gk0, gk1, gk2, ysize, stp100;
	const float *ht, *rocp, *ro, *tk0, *et, *vx, *vy, *vz, *se;
	const double *pr;
};

__global__ void k_heat_source(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
			      const int *__restrict__ bp, HeatSrc S, float *src, double *part)
{
	__shared__ double sh[NB / 64];
	double mx = 0;
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 < nx) {
		long sx = ny, sz = (long)nx * ny, N = sz * nz;
		long v3 = m3 * sz + m1 * sx + m2;
		if (bp[v3] == -2) {
			double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
			double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
			double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
			double rocpkf = (double)S.rocp[v3];
			double ival = 0, x = 0, y = 0, z = 0, v[3] = {0, 0, 0};
			ival += (double)S.ht[v3];
			if (S.adiabyn == 2 || S.frictyn) {
				x = (gx[m1] - gx[m1 - 1]) / 2.0 * xkf;
				y = (gy[m2] - gy[m2 - 1]) / 2.0 * ykf;
				z = (gz[m3] - gz[m3 - 1]) / 2.0 * zkf;
			}
			if (S.adiabyn) vxyz(S.vx, S.vy, S.vz, gx, gy, gz, nx, ny, nz, gx[m1], gy[m2], gz[m3], v);
			bool act = S.ro[v3] > 2e+3f && gy[m2] < S.ysize - S.stp100;
			/* cell-centre corners, y fastest: c0 = v3, c1 = +y, c2 = +x,
			 * c3 = +x+y, c4 = +z, ... c7 = +x+y+z */
			long c0 = v3, c1 = v3 + 1, c2 = v3 + sx, c3 = c2 + 1;
			long c4 = v3 + sz, c5 = c4 + 1, c6 = c4 + sx, c7 = c6 + 1;
			if (S.adiabyn == 2 && act) {
				const double *pr = S.pr;
				double dpdx = xkf * ((1.0 - y) * (1.0 - z) * (pr[c2] - pr[c0]) + y * (1.0 - z) * (pr[c3] - pr[c1]) +
						     (1.0 - y) * z * (pr[c6] - pr[c4]) + y * z * (pr[c7] - pr[c5]));
				double dpdy = ykf * ((1.0 - x) * (1.0 - z) * (pr[c1] - pr[c0]) + x * (1.0 - z) * (pr[c3] - pr[c2]) +
						     (1.0 - x) * z * (pr[c5] - pr[c4]) + x * z * (pr[c7] - pr[c6]));
				double dpdz = zkf * ((1.0 - x) * (1.0 - y) * (pr[c4] - pr[c0]) + x * (1.0 - y) * (pr[c6] - pr[c2]) +
						     (1.0 - x) * y * (pr[c5] - pr[c1]) + x * y * (pr[c7] - pr[c3]));
				float at = S.et[v3] * S.tk0[v3];
				ival += (double)at * (v[0] * dpdx + v[1] * dpdy + v[2] * dpdz);
			}
			if (S.adiabyn == 1 && act) {
				float atr = S.et[v3] * S.tk0[v3] * S.ro[v3];
				ival += (double)atr * (v[0] * S.gk0 + v[1] * S.gk1 + v[2] * S.gk2);
			}
			if (S.frictyn && act) {
				const float *sxx = S.se, *syy = sxx + N, *szz = syy + N, *sxy = szz + N, *sxz = sxy + N,
					    *syz = sxz + N, *exx = syz + N, *eyy = exx + N, *ezz = eyy + N, *exy = ezz + N,
					    *exz = exy + N, *eyz = exz + N;
				long cc[8] = {c0, c1, c2, c3, c4, c5, c6, c7};
				double e = 0;
#pragma unroll
				for (int q = 0; q < 8; q++) {
					long u = cc[q];
					float sn = sxx[u] * exx[u] + syy[u] * eyy[u] + szz[u] * ezz[u];
					double wx = (q >> 1) & 1 ? x : 1.0 - x, wy = q & 1 ? y : 1.0 - y,
					       wz = (q >> 2) & 1 ? z : 1.0 - z;
					e += wx * wy * wz * (double)sn;
				}
				e += 2.0 * ((1.0 - z) * (double)(sxy[v3 - sz] * exy[v3 - sz]) + z * (double)(sxy[v3] * exy[v3]));
				e += 2.0 * ((1.0 - y) * (double)(sxz[v3 - 1] * exz[v3 - 1]) + y * (double)(sxz[v3] * exz[v3]));
				e += 2.0 * ((1.0 - x) * (double)(syz[v3 - sx] * eyz[v3 - sx]) + x * (double)(syz[v3] * eyz[v3]));
				ival += e;
			}
			src[v3] = (float)ival;
			mx = fabs(ival / rocpkf);
		} else {
			src[v3] = 0.0f;
		}
	}
	for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) sh[w] = mx;
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0;
		for (int k = 0; k < NB / 64; k++) a = fmax(a, sh[k]);
		part[(long)blockIdx.y * gridDim.x + blockIdx.x] = a;
	}
}

/* rocp = ro * cp (float, as the reference's product of two float fields),
 * cp arriving in rocp */
__global__ void k_rocp(long N, const float *__restrict__ ro, float *__restrict__ rocp)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < N) rocp[t] = ro[t] * rocp[t];
}

/* block maxima of |H / (ro cp)| over the non-boundary nodes (the step
 * limit of heat3mg.c:327-368 without adiabatic and shear terms) */
__global__ void __launch_bounds__(NB) k_rate_part(int nx, int ny, int nz, const float *__restrict__ ht,
						  const float *__restrict__ rocp, const int *__restrict__ bp, double *part)
{
	__shared__ double sh[NB / 64];
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	double mx = 0;
	if (m1 < nx) {
		long v3 = (long)m3 * nx * ny + (long)m1 * ny + m2;
		if (bp[v3] == -2) mx = fabs((double)ht[v3] / (double)rocp[v3]);
	}
	for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) sh[w] = mx;
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0;
		for (int k = 0; k < NB / 64; k++) a = fmax(a, sh[k]);
		part[(long)blockIdx.y * gridDim.x + blockIdx.x] = a;
	}
}

__global__ void k_heat_max(long nblk, const double *part, double *res)
{
	__shared__ double sh[NB];
	double a = 0;
	for (long i = threadIdx.x; i < nblk; i += NB) a = fmax(a, part[i]);
	sh[threadIdx.x] = a;
	__syncthreads();
	for (int o = NB / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0) res[0] = sh[0];
}

/* Right-hand side and diagonal of every row of one sub-step; boundary rows
 * take their constant, on the bottom line replaced by the plume ramp of
 * titerate (heat3mg.c:444-455).  src: the heat source as tk1 holds it. */
__global__ void k_heat_setup(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
			     const float *__restrict__ kt, const float *__restrict__ rocp,
			     const float *__restrict__ src, const float *__restrict__ tk0, const int *__restrict__ bp,
			     const float *__restrict__ bc, double dt, int plume, double timebeg, double timeend,
			     double timesum, double tempend, double *b, double *diag, double *T)
{
	NODE_IDX
	T[v3] = (double)tk0[v3];
	if (bp[v3] != -2) {
		double c = bc[v3];
		if (plume && m2 == ny - 1)
			c = plume == 2 ? tempend : c + (tempend - c) / (timeend - timebeg) * (timesum - timebeg);
		b[v3] = c;
		diag[v3] = 1.0;
		return;
	}
	double rc = (double)rocp[v3];
	Cond c = conds(kt, gx, gy, gz, m1, m2, m3, v3, sx, sz);
	b[v3] = (double)src[v3] + (double)tk0[v3] / dt * rc;
	double d = rc / dt;
	d += c.xm; d += c.xp; d += c.ym; d += c.yp; d += c.zm; d += c.zp;
	diag[v3] = d;
}

/* One colour of the red-black relaxation of the interior rows:
 * T -= (A T - b) / diag * t0koef (gausmatt, heat3mg.c:933-940).
 * Threads cover only the colour's nodes (every other y index of a line);
 * the boundary code is read first and then all 17 operands of the row are
 * loaded before any is used (the scheduling barrier keeps them together);
 * the arithmetic is conds() and the update term for term.  Work-group
 * g of the 1-D grid (8 * ceil(chunks * nz / 8) groups, handed to the 8
 * XCDs round-robin) takes item (g % 8) * per + g / 8, so every XCD sweeps a
 * contiguous slab of planes and finds the neighbouring planes in its L2. */
__global__ void __launch_bounds__(NB)
k_heat_sweep(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
	     const float *__restrict__ kt, const int *__restrict__ bp, const double *__restrict__ b,
	     const double *__restrict__ diag, double *T, int colour, double t0koef, int chunks)
{
	int per = gridDim.x >> 3;
	int w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
	if (w >= chunks * nz) return;
	int m3 = w / chunks, ch = w - m3 * chunks;
	int t = ch * blockDim.x + threadIdx.x;
	int hy = (ny + 1) >> 1;
	int m1 = t / hy, tj = t - m1 * hy;
	if (m1 >= nx) return;
	int m2 = 2 * tj + ((m1 + m3 + colour) & 1);
	if (m2 >= ny) return;
	long sx = ny, sz = (long)nx * ny;
	long v3 = m3 * sz + m1 * sx + m2;
	if (bp[v3] != -2) return; /* boundary rows (every outer node) */
	float k3 = kt[v3], kxm = kt[v3 - sx], kxp = kt[v3 + sx], kym = kt[v3 - 1], kyp = kt[v3 + 1];
	float kzm = kt[v3 - sz], kzp = kt[v3 + sz];
	double t0 = T[v3], txm = T[v3 - sx], txp = T[v3 + sx], tym = T[v3 - 1], typ = T[v3 + 1];
	double tzm = T[v3 - sz], tzp = T[v3 + sz];
	double dg = diag[v3], bb = b[v3];
	double gx0 = gx[m1 - 1], gx1 = gx[m1], gx2 = gx[m1 + 1];
	double gy0 = gy[m2 - 1], gy1 = gy[m2], gy2 = gy[m2 + 1];
	double gz0 = gz[m3 - 1], gz1 = gz[m3], gz2 = gz[m3 + 1];
	__builtin_amdgcn_sched_barrier(0);
	double xkf = 2.0 / (gx2 - gx0), ykf = 2.0 / (gy2 - gy0), zkf = 2.0 / (gz2 - gz0);
	double cxm = (double)(kxm + k3) / 2.0 * xkf / (gx1 - gx0);
	double cxp = (double)(k3 + kxp) / 2.0 * xkf / (gx2 - gx1);
	double cym = (double)(kym + k3) / 2.0 * ykf / (gy1 - gy0);
	double cyp = (double)(k3 + kyp) / 2.0 * ykf / (gy2 - gy1);
	double czm = (double)(kzm + k3) / 2.0 * zkf / (gz1 - gz0);
	double czp = (double)(k3 + kzp) / 2.0 * zkf / (gz2 - gz1);
	double r = dg * t0 - bb;
	r -= cxm * txm; r -= cxp * txp;
	r -= cym * tym; r -= cyp * typ;
	r -= czm * tzm; r -= czp * tzp;
	T[v3] = t0 - r / dg * t0koef;
}

/* Boundary rows of one dependency depth: T = c + k T[partner]. */
__global__ void k_heat_bc(long n, const int *__restrict__ list, const int *__restrict__ bp,
			  const float *__restrict__ bk, const double *__restrict__ b, double *T, double t0koef)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int v = list[i];
	int p = bp[v];
	double r = T[v] - b[v];
	if (p >= 0) r -= (double)bk[v] * T[p];
	T[v] -= r * t0koef;
}

/* Sums of the squared heat-row errors in temperature units (heaterr
 * ynerr=2 times the step, heat3mg.c:512-537) and of the boundary-row errors
 * (tbonderr ynerr=1), per block. */
__global__ void k_heat_norm(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
			    const float *__restrict__ kt, const float *__restrict__ rocp, const int *__restrict__ bp,
			    const float *__restrict__ bk, const double *__restrict__ b,
			    const double *__restrict__ diag, const double *__restrict__ T, double dt, double *part)
{
	__shared__ double sh[2][NB / 64];
	double h = 0, q = 0;
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 < nx) {
		long sx = ny, sz = (long)nx * ny;
		long v3 = m3 * sz + m1 * sx + m2;
		int p = bp[v3];
		if (p != -2) {
			double r = T[v3] - b[v3];
			if (p >= 0) r -= (double)bk[v3] * T[p];
			q = r * r;
		} else {
			Cond c = conds(kt, gx, gy, gz, m1, m2, m3, v3, sx, sz);
			double r = diag[v3] * T[v3] - b[v3];
			r -= c.xm * T[v3 - sx]; r -= c.xp * T[v3 + sx];
			r -= c.ym * T[v3 - 1];  r -= c.yp * T[v3 + 1];
			r -= c.zm * T[v3 - sz]; r -= c.zp * T[v3 + sz];
			r = r / (double)rocp[v3] * dt;
			h = r * r;
		}
	}
	for (int o = 32; o > 0; o >>= 1) {
		h += __shfl_down(h, o, 64);
		q += __shfl_down(q, o, 64);
	}
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) { sh[0][w] = h; sh[1][w] = q; }
	__syncthreads();
	if (threadIdx.x == 0) {
		double a = 0, c = 0;
		for (int k = 0; k < NB / 64; k++) { a += sh[0][k]; c += sh[1][k]; }
		long blk = (long)blockIdx.y * gridDim.x + blockIdx.x;
		part[2 * blk] = a;
		part[2 * blk + 1] = c;
	}
}

__global__ void k_heat_sum(long nblk, const double *part, double *res)
{
	__shared__ double sh[2][NB];
	double a = 0, c = 0;
	for (long i = threadIdx.x; i < nblk; i += NB) {
		a += part[2 * i];
		c += part[2 * i + 1];
	}
	sh[0][threadIdx.x] = a;
	sh[1][threadIdx.x] = c;
	__syncthreads();
	for (int o = NB / 2; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) {
			sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
			sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) { res[0] = sh[0][0]; res[1] = sh[1][0]; }
}

/* end of a sub-step: the reference stores T as float (tk[] = sol0[]) and
 * starts the next sub-step from that float field */
__global__ void k_heat_store(long N, double *T, float *tk0)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N) return;
	float f = (float)T[i];
	tk0[i] = f;
	T[i] = f;
}

/* ------------------------------------------------------------------ */
/* C ABI                                                              */
/* ------------------------------------------------------------------ */

extern "C" int i3g_heat_destroy(i3g_heat *h)
{
	if (!h) return I3G_OK;
	if (h->s) hipStreamSynchronize(h->s);
	void *ps[] = {h->gx, h->gy, h->gz, h->bp, h->bc, h->bk, h->blist, h->kt, h->rocp, h->ht, h->tk0,
		      h->T, h->b, h->diag, h->part, h->res, h->vx, h->vy, h->vz, h->et, h->ro, h->se, h->pr, h->src};
	for (void *p : ps)
		if (p) hipFree(p);
	if (h->s) hipStreamDestroy(h->s);
	delete h;
	return I3G_OK;
}

extern "C" int i3g_heat_create(const i3g_setup *su, i3g_heat **out)
{
	if (!su || !out || !su->gx || !su->gy || !su->gz || !su->bondm || !su->bondv || !su->bondn)
		return I3G_EINVAL;
	if (su->nx < 3 || su->ny < 3 || su->nz < 3 || su->nrec < 1) return I3G_EINVAL;
	long N = su->nx * su->ny * su->nz;
	if (N > 400000000L) return I3G_ECAP;
	*out = nullptr;
	i3g_heat *h = new i3g_heat();
	h->nx = (int)su->nx; h->ny = (int)su->ny; h->nz = (int)su->nz; h->N = N;
	/* boundary rows of the T block (bondm + 8*nodenum) */
	std::vector<int> bp(N, -2);
	std::vector<float> bc(N, 0.f), bk(N, 0.f);
	const int *bm = su->bondm + 8 * N;
	for (long i = 0; i < N; i++) {
		long r = bm[i];
		if (r <= 0) continue;
		if (r >= su->nrec) { delete h; return I3G_EBC; }
		long p = (long)su->bondn[r] - 1;
		if (p < -1 || p >= N) { delete h; return I3G_EBC; }
		bp[i] = (int)p;
		bc[i] = su->bondv[2 * r];
		bk[i] = su->bondv[2 * r + 1];
	}
	/* a boundary node on the domain surface must exist for every surface
	 * node: the interior stencil reads all six neighbours */
	for (long m3 = 0; m3 < su->nz; m3++)
		for (long m1 = 0; m1 < su->nx; m1++)
			for (long m2 = 0; m2 < su->ny; m2++) {
				bool surf = m1 == 0 || m2 == 0 || m3 == 0 || m1 == su->nx - 1 || m2 == su->ny - 1 ||
					    m3 == su->nz - 1;
				long v = m3 * su->nx * su->ny + m1 * su->ny + m2;
				if (surf && bp[v] == -2) { delete h; return I3G_EBC; }
			}
	/* dependency depth of the boundary rows */
	std::vector<int> depth(N, -1);
	std::vector<std::vector<int>> lists(MAXDEPTH);
	for (long i = 0; i < N; i++) {
		if (bp[i] == -2) continue;
		int d = 0;
		long j = i;
		while (bp[j] >= 0 && bp[bp[j]] != -2) {
			j = bp[j];
			if (++d >= MAXDEPTH) { delete h; return I3G_EBC; }
		}
		depth[i] = d;
		lists[d].push_back((int)i);
	}
	std::vector<int> flat;
	h->boff.push_back(0);
	for (int d = 0; d < MAXDEPTH; d++) {
		flat.insert(flat.end(), lists[d].begin(), lists[d].end());
		h->boff.push_back((long)flat.size());
	}
	int err = 0;
	if (hipStreamCreate(&h->s) != hipSuccess) { delete h; return I3G_EDEVICE; }
	h->gx = halloc<double>(su->nx, &err); h->gy = halloc<double>(su->ny, &err); h->gz = halloc<double>(su->nz, &err);
	h->bp = halloc<int>(N, &err); h->bc = halloc<float>(N, &err); h->bk = halloc<float>(N, &err);
	h->blist = halloc<int>(flat.size(), &err);
	h->kt = halloc<float>(N, &err); h->rocp = halloc<float>(N, &err); h->ht = halloc<float>(N, &err);
	h->ro = halloc<float>(N, &err);
	h->tk0 = halloc<float>(N, &err);
	h->T = halloc<double>(N, &err); h->b = halloc<double>(N, &err); h->diag = halloc<double>(N, &err);
	long nblk = ((long)su->nx * su->ny + NB - 1) / NB * su->nz;
	h->part = halloc<double>(2 * nblk, &err);
	h->res = halloc<double>(2, &err);
	if (err) { i3g_heat_destroy(h); return err; }
	if (hipMemcpy(h->gx, su->gx, su->nx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->gy, su->gy, su->ny * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->gz, su->gz, su->nz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->bp, bp.data(), N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->bc, bc.data(), N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(h->bk, bk.data(), N * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
	    (flat.size() &&
	     hipMemcpy(h->blist, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
		i3g_heat_destroy(h);
		return I3G_EDEVICE;
	}
	h->hb.assign(bm, bm + N);
	*out = h;
	return I3G_OK;
}

extern "C" int i3g_heat_default_params(i3g_heat_params *p)
{
	if (!p) return I3G_EINVAL;
	p->timestep = 0;
	p->maxtkstep = 25.0;  /* mode.t3c maxtkstep (K) */
	p->heatmin = 1e-3;    /* mode.t3c HEATMIN */
	p->t0koef = 1.0;
	p->sweeps = 30;       /* multittt[0][0] */
	p->max_sweeps = 100000;
	p->timesum = 0;
	p->timebeg = 1e300;
	p->timeend = 1e300;
	p->tempend = 0;
	p->verbose = 0;
	return I3G_OK;
}

static int bc_apply(i3g_heat *h, double t0koef)
{
	for (size_t d = 0; d + 1 < h->boff.size(); d++) {
		long n = h->boff[d + 1] - h->boff[d];
		if (n <= 0) continue;
		hipLaunchKernelGGL(k_heat_bc, dim3((unsigned)((n + NB - 1) / NB)), dim3(NB), 0, h->s, n,
				   (const int *)(h->blist + h->boff[d]), (const int *)h->bp, (const float *)h->bk,
				   (const double *)h->b, h->T, t0koef);
	}
	return I3G_OK;
}

extern "C" int i3g_heat_set_sources(i3g_heat *h, int adiabyn, int frictyn, const double *gk, double ysize,
				    double stp100, const float *vx, const float *vy, const float *vz, const double *pr,
				    const float *et, const float *const *sig, const float *const *eps)
{
	if (!h || adiabyn < 0 || adiabyn > 2 || frictyn < 0 || frictyn > 1) return I3G_EINVAL;
	h->adiabyn = adiabyn;
	h->frictyn = frictyn;
	if (!adiabyn && !frictyn) return I3G_OK;
	if (!gk || !et || (adiabyn && (!vx || !vy || !vz)) || (adiabyn == 2 && !pr) || (frictyn && (!sig || !eps))) {
		h->adiabyn = h->frictyn = 0;
		return I3G_EINVAL;
	}
	long N = h->N;
	int err = 0;
	if (!h->vx) {
		h->vx = halloc<float>(N, &err); h->vy = halloc<float>(N, &err); h->vz = halloc<float>(N, &err);
		h->et = halloc<float>(N, &err); h->se = halloc<float>(12 * N, &err);
		h->pr = halloc<double>(N, &err); h->src = halloc<float>(N, &err);
		if (err) { h->adiabyn = h->frictyn = 0; return err; }
	}
	for (int q = 0; q < 3; q++) h->gk[q] = gk[q];
	h->ysize = ysize;
	h->stp100 = stp100;
	size_t fb = N * sizeof(float);
	TCHK(hipMemcpyAsync(h->et, et, fb, hipMemcpyHostToDevice, h->s));
	if (adiabyn) {
		TCHK(hipMemcpyAsync(h->vx, vx, fb, hipMemcpyHostToDevice, h->s));
		TCHK(hipMemcpyAsync(h->vy, vy, fb, hipMemcpyHostToDevice, h->s));
		TCHK(hipMemcpyAsync(h->vz, vz, fb, hipMemcpyHostToDevice, h->s));
	}
	if (adiabyn == 2) TCHK(hipMemcpyAsync(h->pr, pr, N * sizeof(double), hipMemcpyHostToDevice, h->s));
	if (frictyn)
		for (int q = 0; q < 6; q++) {
			if (!sig[q] || !eps[q]) { h->adiabyn = h->frictyn = 0; return I3G_EINVAL; }
			TCHK(hipMemcpyAsync(h->se + q * N, sig[q], fb, hipMemcpyHostToDevice, h->s));
			TCHK(hipMemcpyAsync(h->se + (6 + q) * N, eps[q], fb, hipMemcpyHostToDevice, h->s));
		}
	TCHK(hipStreamSynchronize(h->s));
	return I3G_OK;
}

extern "C" int i3g_heat_solve(i3g_heat *h, float *tk, const float *ro, const float *cp, const float *kt,
			      const float *ht, const i3g_heat_params *p, i3g_heat_stats *st)
{
	if (!h || !tk || !ro || !cp || !kt || !ht || !p) return I3G_EINVAL;
	if (!(p->t0koef > 0 && p->t0koef < 2) || p->sweeps < 1 || p->max_sweeps < 1) return I3G_EINVAL;
	auto t_start = std::chrono::steady_clock::now();
	long N = h->N;
	i3g_heat_stats s{};
	/* temperature step limit (heat3mg.c:327-368) from max |src/(ro cp)|
	 * over the non-boundary nodes; src = H, or computed on the device with
	 * the adiabatic/shear terms.  The fields may be host or device arrays. */
	bool terms = h->adiabyn || h->frictyn;
	dim3 gg((unsigned)(((long)h->nx * h->ny + NB - 1) / NB), (unsigned)h->nz);
	long nblk = (long)gg.x * gg.y;
	/* sweep grid: one thread per node of a colour, XCD-contiguous planes */
	int swchunks = (int)(((long)h->nx * ((h->ny + 1) / 2) + NB - 1) / NB);
	dim3 gsw((unsigned)(8 * (((long)swchunks * h->nz + 7) / 8)));
	double maxrate = 0;
	TCHK(hipMemcpyAsync(h->rocp, cp, N * sizeof(float), hipMemcpyDefault, h->s));
	TCHK(hipMemcpyAsync(h->ro, ro, N * sizeof(float), hipMemcpyDefault, h->s));
	TCHK(hipMemcpyAsync(h->ht, ht, N * sizeof(float), hipMemcpyDefault, h->s));
	TCHK(hipMemcpyAsync(h->tk0, tk, N * sizeof(float), hipMemcpyDefault, h->s));
	TCHK(hipMemcpyAsync(h->kt, kt, N * sizeof(float), hipMemcpyDefault, h->s));
	hipLaunchKernelGGL(k_rocp, dim3((unsigned)((N + NB - 1) / NB)), dim3(NB), 0, h->s, N, (const float *)h->ro, h->rocp);
	if (terms) {
		HeatSrc S{h->adiabyn, h->frictyn, h->gk[0], h->gk[1], h->gk[2], h->ysize, h->stp100, h->ht, h->rocp,
			  h->ro, h->tk0, h->et, h->vx, h->vy, h->vz, h->se, h->pr};
		hipLaunchKernelGGL(k_heat_source, gg, dim3(NB), 0, h->s, h->nx, h->ny, h->nz, (const double *)h->gx,
				   (const double *)h->gy, (const double *)h->gz, (const int *)h->bp, S, h->src, h->part);
	} else {
		hipLaunchKernelGGL(k_rate_part, gg, dim3(NB), 0, h->s, h->nx, h->ny, h->nz, (const float *)h->ht,
				   (const float *)h->rocp, (const int *)h->bp, h->part);
	}
	hipLaunchKernelGGL(k_heat_max, dim3(1), dim3(NB), 0, h->s, nblk, (const double *)h->part, h->res);
	TCHK(hipGetLastError());
	TCHK(hipMemcpyAsync(&maxrate, h->res, sizeof(double), hipMemcpyDeviceToHost, h->s));
	TCHK(hipStreamSynchronize(h->s));
	if (p->timestep <= 0) {
		/* the reference returns without a solve when the step is zero */
		if (st) *st = s;
		return p->timestep == 0 ? I3G_OK : I3G_EINVAL;
	}
	double tkstep = 1e30;
	if (maxrate > 0 && p->maxtkstep > 0) tkstep = p->maxtkstep / maxrate;
	if (maxrate > 0 && p->maxtkstep == 0) {
		if (st) *st = s;
		return I3G_OK;   /* tktimestep == 0: the reference returns */
	}
	int plume = p->timesum > p->timebeg ? (p->timesum >= p->timeend ? 2 : 1) : 0;
	hipEvent_t e0, e1;
	TCHK(hipEventCreate(&e0));
	TCHK(hipEventCreate(&e1));
	TCHK(hipEventRecord(e0, h->s));
	double total = p->timestep, done = 0;
	double hres[2] = {0, 0};
	int rc = I3G_OK;
	while (done < total) {
		double dt = total < tkstep ? total : tkstep;
		if (dt > total - done) dt = total - done;
		hipLaunchKernelGGL(k_heat_setup, gg, dim3(NB), 0, h->s, h->nx, h->ny, h->nz, (const double *)h->gx,
				   (const double *)h->gy, (const double *)h->gz, (const float *)h->kt,
				   (const float *)h->rocp, (const float *)(terms ? h->src : h->ht), (const float *)h->tk0,
				   (const int *)h->bp, (const float *)h->bc, dt, plume, p->timebeg, p->timeend, p->timesum,
				   p->tempend, h->b, h->diag,
				   h->T);
		int sw = 0;
		for (;;) {
			for (int k = 0; k < p->sweeps; k++) {
				for (int col = 0; col < 2; col++) {
					hipLaunchKernelGGL(k_heat_sweep, gsw, dim3(NB), 0, h->s, h->nx, h->ny, h->nz,
							   (const double *)h->gx, (const double *)h->gy,
							   (const double *)h->gz, (const float *)h->kt, (const int *)h->bp,
							   (const double *)h->b, (const double *)h->diag, h->T, col,
							   p->t0koef, swchunks);
					bc_apply(h, p->t0koef);
				}
			}
			sw += p->sweeps;
			hipLaunchKernelGGL(k_heat_norm, gg, dim3(NB), 0, h->s, h->nx, h->ny, h->nz, (const double *)h->gx,
					   (const double *)h->gy, (const double *)h->gz, (const float *)h->kt,
					   (const float *)h->rocp, (const int *)h->bp, (const float *)h->bk,
					   (const double *)h->b, (const double *)h->diag, (const double *)h->T, dt, h->part);
			hipLaunchKernelGGL(k_heat_sum, dim3(1), dim3(NB), 0, h->s, nblk, (const double *)h->part, h->res);
			TCHK(hipGetLastError());
			TCHK(hipMemcpyAsync(hres, h->res, 2 * sizeof(double), hipMemcpyDeviceToHost, h->s));
			TCHK(hipStreamSynchronize(h->s));
			long nb = h->boff.back();
			s.heatsum = std::sqrt(hres[0] / (double)(N - nb));
			s.bondsum = nb ? std::sqrt(hres[1] / (double)nb) : 0;
			if (p->verbose >= 2)
				fprintf(stderr, "i3g heat: sub-step %d sweeps %d heatsum %.3e bondsum %.3e\n", s.substeps + 1,
					sw, s.heatsum, s.bondsum);
			if (!(s.heatsum > p->heatmin)) break;
			if (sw >= p->max_sweeps) { rc = I3G_ENOCONV; break; }
		}
		s.sweeps += sw;
		hipLaunchKernelGGL(k_heat_store, dim3((unsigned)((N + NB - 1) / NB)), dim3(NB), 0, h->s, N, h->T, h->tk0);
		s.substeps++;
		done += dt;
		if (rc) break;
	}
	TCHK(hipGetLastError());
	TCHK(hipEventRecord(e1, h->s));
	TCHK(hipMemcpyAsync(tk, h->tk0, N * sizeof(float), hipMemcpyDefault, h->s));
	TCHK(hipStreamSynchronize(h->s));
	{
		float ms = 0;
		if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) s.device_seconds = ms * 1e-3;
		hipEventDestroy(e0);
		hipEventDestroy(e1);
	}
	s.converged = rc == I3G_OK;
	s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
	if (st) *st = s;
	return rc;
}
