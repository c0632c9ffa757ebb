/*
 * i3g Stokes path on MI355X (gfx950): device-resident multigrid for the
 * variable-viscosity staggered-grid Stokes system of the reference
 * (move3mg.c), matrix-free, with an FP64 outer Krylov/defect-correction
 * loop and FP64 or FP32 V-cycles.  C ABI in include/i3g.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/i3g.h"
#include "stokes_kernels.h"

#define HCHK(x)                                                                \
	do {                                                                   \
		hipError_t e_ = (x);                                           \
		if (e_ != hipSuccess) {                                        \
			fprintf(stderr, "i3g: HIP error %s at %s:%d\n",       \
				hipGetErrorString(e_), __FILE__, __LINE__);    \
			return I3G_EDEVICE;                                    \
		}                                                              \
	} while (0)

/* ------------------------------------------------------------------ */
/* Coefficient kernels                                                */
/* ------------------------------------------------------------------ */

/* Effective viscosities at the stress points of one level (the nueff
 * blocks of sxxcalc/continerr move3mg.c:36-48, sxycalc :402-414, sxzcalc
 * :517-529, syzcalc :632-644).  viscmod 0: the reference adds the float
 * node values in float before dividing in double. */
__global__ void k_eta(int nx, int ny, int nz, int lev0, const float *__restrict__ nu,
		      const float *__restrict__ nxx, const float *__restrict__ nxy,
		      const float *__restrict__ nxz, const float *__restrict__ nyz,
		      int viscmod, double nubeg, double nuend, double *e64, float *e32,
		      double *wP64, float *wP32, const double *gx, const double *gy,
		      const double *gz)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 >= nx) return;
	long N = (long)nx * ny * nz, sx = ny, sz = (long)nx * ny;
	long n = m3 * sz + m1 * sx + m2;
	double eC = 0, eXY = 0, eXZ = 0, eYZ = 0, e;
	if (m1 >= 1 && m2 >= 1 && m3 >= 1) {
		if (lev0 && m1 > 3 && m2 > 3 && m3 > 3 && m1 < nx - 3 && m2 < ny - 3 && m3 < nz - 3) {
			e = nxx[n];
		} else {
			long v0 = n - sz - sx - 1;
			long v[8] = {v0, v0 + 1, v0 + sx, v0 + sx + 1, v0 + sz, v0 + sz + 1, v0 + sz + sx, v0 + sz + sx + 1};
			if (viscmod == 1) {
				double s = 0;
				for (int q = 0; q < 8; q++) s += log((double)nu[v[q]]);
				e = exp(s / 8.0);
			} else if (viscmod == 2) {
				double s = 0;
				for (int q = 0; q < 8; q++) s += 1.0 / (double)nu[v[q]];
				e = 1.0 / (s / 8.0);
			} else {
				float fs = nu[v[0]];
				for (int q = 1; q < 8; q++) fs = fs + nu[v[q]];
				e = (double)fs / 8.0;
			}
		}
		eC = e > nuend ? nuend : (e < nubeg ? nubeg : e);
	}
#define AVG2(a, b)                                                                          \
	(viscmod == 1 ? exp((log((double)(a)) + log((double)(b))) / 2.0)                    \
	 : viscmod == 2 ? 1.0 / ((1.0 / (double)(a) + 1.0 / (double)(b)) / 2.0)           \
			: (double)(float)((a) + (b)) / 2.0)
	if (m3 < nz - 1) {
		if (lev0 && m1 > 3 && m2 > 3 && m3 > 2 && m1 < nx - 4 && m2 < ny - 4 && m3 < nz - 4)
			e = nxy[n];
		else
			e = AVG2(nu[n], nu[n + sz]);
		eXY = e > nuend ? nuend : (e < nubeg ? nubeg : e);
	}
	if (m2 < ny - 1) {
		if (lev0 && m1 > 3 && m2 > 2 && m3 > 3 && m1 < nx - 4 && m2 < ny - 4 && m3 < nz - 4)
			e = nxz[n];
		else
			e = AVG2(nu[n], nu[n + 1]);
		eXZ = e > nuend ? nuend : (e < nubeg ? nubeg : e);
	}
	if (m1 < nx - 1) {
		if (lev0 && m1 > 2 && m2 > 3 && m3 > 3 && m1 < nx - 4 && m2 < ny - 4 && m3 < nz - 4)
			e = nyz[n];
		else
			e = AVG2(nu[n], nu[n + sx]);
		eYZ = e > nuend ? nuend : (e < nubeg ? nubeg : e);
	}
#undef AVG2
	e64[n] = eC; e64[N + n] = eXY; e64[2 * N + n] = eXZ; e64[3 * N + n] = eYZ;
	e32[n] = (float)eC; e32[N + n] = (float)eXY; e32[2 * N + n] = (float)eXZ; e32[3 * N + n] = (float)eYZ;
	double w = 0;
	if (m1 >= 1 && m2 >= 1 && m3 >= 1) {
		double h = ((gx[m1] - gx[m1 - 1]) + (gy[m2] - gy[m2 - 1]) + (gz[m3] - gz[m3 - 1])) / 3.0;
		w = eC / h;
	}
	wP64[n] = w; wP32[n] = (float)w;
}

/* Momentum right-hand side -g_i * mean of four node densities, summed in
 * float as the reference (move3mg.c:751, 943, 1142); BC rows take their
 * constant; continuity rows 0. */
__global__ void k_rhs(int nx, int ny, int nz, const float *__restrict__ ro, double g0,
		      double g1, double g2, const uint8_t *__restrict__ flags,
		      const int *__restrict__ bcid, const double *__restrict__ bcc,
		      double *__restrict__ b)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int k = blockIdx.y;
	int i = t / ny, j = t - i * ny;
	if (i >= nx) return;
	long N = (long)nx * ny * nz, sx = ny, sz = (long)nx * ny;
	long n = k * sz + i * sx + j;
	unsigned fl = flags[n];
	for (int c = 0; c < 4; c++) {
		double v = 0;
		long u = c * N + n;
		if (fl & FL_EX(c)) {
			if (fl & FL_BC(c)) v = bcc[bcid[u]];
			else if (c == 1) v = -g0 * (ro[n] + ro[n + 1] + ro[n + sz] + ro[n + 1 + sz]) / 4.0;
			else if (c == 2) v = -g1 * (ro[n] + ro[n + sx] + ro[n + sz] + ro[n + sx + sz]) / 4.0;
			else if (c == 3) v = -g2 * (ro[n] + ro[n + 1] + ro[n + sx] + ro[n + sx + 1]) / 4.0;
		}
		b[u] = v;
	}
}

/* Coarse viscosity restriction in gather form (ronurestrict,
 * move3mg.c:2642-2741): float accumulator, double weight sum, fine nodes
 * in ascending index order as the reference's serial scatter. */
__global__ void k_ronurestrict(int fnx, int fny, int cnx, int cny, int cnz,
			       const float *__restrict__ nuf, float *__restrict__ nuc,
			       const int *xs, const int *xi, const double *xw,
			       const int *ys, const int *yi, const double *yw,
			       const int *zs, const int *zi, const double *zw, int mode)
{
#pragma clang fp contract(off)
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int K = blockIdx.y;
	int I = t / cny, J = t - I * cny;
	if (I >= cnx) return;
	long szf = (long)fnx * fny;
	float acc = 0;
	double sw = 0;
	for (int a = zs[K]; a < zs[K + 1]; a++)
		for (int b = xs[I]; b < xs[I + 1]; b++)
			for (int e = ys[J]; e < ys[J + 1]; e++) {
				double swt = xw[b] * yw[e] * zw[a];
				long nf = zi[a] * szf + (long)xi[b] * fny + yi[e];
				/* viscmod 0/1/2: the value, its log or its inverse */
				double term = mode == 0 ? (double)nuf[nf] * swt
					      : mode == 1 ? log((double)nuf[nf]) * swt : 1.0 / (double)nuf[nf] * swt;
				acc = (float)((double)acc + term);
				sw += swt;
			}
	float v = (float)((double)acc / sw);
	if (mode == 1) v = (float)exp((double)v);
	if (mode == 2) v = (float)(1.0 / (double)v);
	nuc[(long)K * cnx * cny + (long)I * cny + J] = v;
}

/* ------------------------------------------------------------------ */
/* Host structures                                                    */
/* ------------------------------------------------------------------ */

template <typename T> static T *dalloc(size_t n, int *err)
{
	void *p = nullptr;
	if (n == 0) n = 1;
	if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
		*err = I3G_ENOMEM;
		return nullptr;
	}
	return (T *)p;
}

template <typename T> static int upload(T *d, const T *h, size_t n)
{
	if (n) HCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
	return I3G_OK;
}

struct AxisMapDev {
	/* kinds 0..2 (vp transfers) and 3 (ronurestrict) */
	int *base[4];
	double *wlo64[4], *whi64[4], *iw64[4];
	float *wlo32[4], *whi32[4], *iw32[4];
	int *istart[4], *iidx[4];
};

struct Level {
	int nx, ny, nz;
	long N;
	double *g[3];            /* lines */
	double *ih64[3], *ic64[3];
	float *ih32[3], *ic32[3];
	double *eta64;           /* 4N: eC, eXY, eXZ, eYZ */
	float *eta32;
	double *wP64;
	float *wP32;
	uint8_t *flags;
	int *bcid;               /* 4N -> record (sorted) */
	int nbc;
	int *bc_u, *bc_p;
	double *bc_c64, *bc_k64;
	float *bc_c32, *bc_k32;
	std::vector<int> depth_start; /* record ranges per depth */
	float *nu;               /* level viscosity at nodes */
	double *x64, *b64, *r64;
	float *x32, *b32, *r32;
	/* second unknown vector of the out-of-place sweeps (k_sweep_comp) and
	 * its tiling */
	double *x64b;
	float *x32b;
	FusedGeom fg;
	std::vector<double> hg[3];
	double *scgs = nullptr;  /* 7N corrections of the coupled cell smoother (I3G_SCGS=1) */
};

struct Xfer {
	AxisMapDev ax[3];
};

struct KStat {
	double ms = 0;
	long calls = 0;
	double bytes = 0;
};

struct i3g_ctx {
	hipStream_t s = nullptr;
	int dev = 0;                 /* HIP device the context lives on */
	i3g_hier h;
	int nlev = 0;
	std::vector<Level> L;
	std::vector<Xfer> X;
	double *rhs0 = nullptr;      /* level-0 FP64 right-hand side */
	double *sol = nullptr;       /* level-0 FP64 solution (4N) */
	float *ro = nullptr, *nxx = nullptr, *nxy = nullptr, *nxz = nullptr, *nyz = nullptr;
	double *partials = nullptr;
	long npart = 0;
	double *dscal = nullptr;
	double *hscal = nullptr;     /* pinned */
	/* Krylov storage */
	/* GCR search directions (p_j, q_j = W A p_j), kept across the outer
	 * refinement steps of one solve */
	int kr_m = 0;
	int kr_prec = 0;
	int kr_k = 0;
	std::vector<void *> P, Q;
	void **dP = nullptr, **dQ = nullptr;
	void *kr = nullptr, *ke = nullptr;
	double *dh = nullptr;
	/* inner FGMRES storage */
	int in_m = 0, in_prec = 0;
	std::vector<void *> iV, iZ;
	void **diV = nullptr, **diZ = nullptr;
	double *e64 = nullptr;
	/* timing */
	bool timing = false;
	std::map<std::string, KStat> stats;
	std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
	std::vector<hipEvent_t> evpool;
	size_t evnext = 0;
	double phys[6] = {0, 9.81, 0, 1e18, 1e25, 0};
	int coef_ready = 0;
	/* coarse part of the V-cycle (levels >= 1) captured as a HIP graph per
	 * precision; key = the parameters it was captured with */
	hipGraphExec_t cg_exec[2] = {nullptr, nullptr};
	hipGraph_t cg_graph[2] = {nullptr, nullptr};
	std::vector<double> cg_key[2];
	int use_graph = 1;
	/* XCD-contiguous plane slabs for the smoother on every level of 64
	 * planes or more (I3G_XCD_REMAP=1: level 0 only, =0 off) */
	int xcd_remap = 2;
	/* viscosity averaging of the coarse (V-cycle only) levels:
	 * -1 the reference's viscmod, or 0 arithmetic, 1 geometric, 2
	 * harmonic (I3G_COARSE_VISC selects) */
	int coarse_visc = -1;
	int coarse_visc_maxlev = 99; /* I3G_COARSE_VISC applies to levels 1..this */
	/* velocity smoother with all row loads issued first (I3G_SMOOTH_ILP=0
	 * selects the plain kernel) */
	int smooth_ilp = 1;
	/* smoother sweeps in z-slabs of this many planes on level 0 (on every
	 * level with slab_all): -1 automatic (equal slabs of at most
	 * SLAB_BYTES of sweep operands), 0 one slab (I3G_SLAB, I3G_SLAB_ALL) */
	int slab = -1;
	int slab_all = 0;
	/* level transfer maps: 0 the reference's index rules, 1 geometric
	 * bracketing (I3G_XFER_GEO=1) */
	int xfer_geo = 0;
	/* coarsest-level sweeps in one work-group from LDS (k_coarsest) when
	 * they fit; I3G_FUSED_BOTTOM=0 switches it off */
	int bottom = 0;
	size_t coarsest_lds_max = 0;
	/* BC rows of all depths in one launch (k_bc_chain); I3G_BC_CHAIN=0
	 * selects one launch per depth */
	int bc_chain = 1;
	/* one launch per velocity component and sweep (k_sweep_comp, out of
	 * place, pairs of sweeps) when I3G_FUSED_SWEEP=1; the default is the
	 * seven phase launches, which measured faster at 257^3 (1.03 against
	 * 1.28 ms a sweep: the component kernels' re-reads of the other
	 * fields miss L2 and their per-tile plane chains hide latency worse).
	 * Tile targets (x, y, z-slab) from I3G_FUSED_TILE=tx,ty,tz */
	int fused = 0;
	int ftile[3] = {16, 64, 16};
	/* whole FP32 sweeps in one launch (k_sweep7) on levels whose planes
	 * split into at least sweep7_min_tiles tiles, when I3G_SWEEP7=1
	 * (I3G_SWEEP7_MIN_TILES sets the threshold).  Off by default: at 257^3
	 * it measured 0.95 ms a sweep against 0.91 ms for the z-slabbed phase
	 * launches (its eight barriers a step and the 1.6x recomputed halo rows
	 * leave it bound by the phases' latency chains, not by memory) */
	int sweep7 = 0;
	int sweep7_min_tiles = 128;
	/* the mixed V-cycle's coarsest level solved in FP64 (I3G_COARSEST64=0:
	 * in the V-cycle's FP32) */
	int coarsest64 = 1;
	/* the coarsest level solved directly in FP64 (I3G_COARSE_DIRECT=1): the
	 * truncated pseudo-inverse of its assembled operator, singular values
	 * below cd_tol times the largest dropped (I3G_COARSE_DIRECT_TOL) */
	int coarse_direct = 0;
	double cd_tol = 1e-10;
	/* The coupled cell smoother (k_scgs_cell / k_scgs_apply) on levels
	 * scgs_minlev..scgs_maxlev (I3G_SCGS=0 off; I3G_SCGS_LEVELS "last" or
	 * "first,last"; damping I3G_SCGS_W "velocity,pressure").  scgs_mix > 0
	 * (I3G_SCGS_MIX): at most scgs_phases colour-phase sweeps
	 * (I3G_SCGS_PHASES) and then scgs_mix coupled sweeps per smoothing;
	 * 0: coupled sweeps only.  Guard (I3G_SCGS_GUARD, V-cycles; 0 off):
	 * when a GCR pass's residual falls by less than 10% (to more than
	 * scgs_guard_ratio of where it was) over that many V-cycles, the
	 * coupled sweeps are switched off for the rest of the context's life
	 * (scgs_off) and the levels get their full colour-phase sweeps.  The defaults are the fastest measured on the 257^3
	 * subduction layout. */
	int scgs = 1, scgs_minlev = 0, scgs_maxlev = 0;
	double scgs_wv = 0.4, scgs_wp = 0.5;
	int scgs_mix = 1, scgs_phases = 6, scgs_guard = 100;
	double scgs_guard_ratio = 0.9; /* I3G_SCGS_GUARD_RATIO */
	int scgs_off = 0;
	int cd_n = 0;                /* unknowns (4 N of the coarsest level) */
	double *cd_u = nullptr;      /* U, column-major n x n */
	double *cd_vt = nullptr;     /* V^T, column-major n x n */
	double *cd_s = nullptr;      /* reciprocal kept singular values (n) */
	double *cd_y = nullptr;      /* work (n) */
	int cd_rank = 0;             /* singular values kept */
	double cd_setup_s = 0;       /* time of the last factorisation */
	/* whole FP32 sweeps in three launches (k_sweep3<1..3>, one per velocity
	 * component, both colours each, P with vz) on levels that give at least
	 * sweep3_min_groups work-groups, when I3G_SWEEP3=1
	 * (I3G_SWEEP3_MIN_GROUPS sets the threshold, I3G_SWEEP3_NZS the
	 * z-slabs).  Off by default: 1.16 ms a sweep at 257^3 against 0.92 ms
	 * for the z-slabbed phase launches; its counters show it bound by
	 * vector-ALU issue (about 500 instructions per wave and plane step:
	 * the ring and prefetch address arithmetic outweighs the relaxation
	 * itself) */
	int sweep3 = 0;
	int sweep3_min_groups = 256;
	int sweep3_nzs = 0;          /* 0: as many slabs as fill two groups per CU */
	size_t sweep3_lds = 0;
	size_t sweep7_lds = 0;       /* LDS opt-in granted to k_sweep7 */
	int fused_minz = 64; /* levels with fewer planes keep the phase launches (I3G_FUSED_MINZ) */
	/* learned start and start combination work buffers (kept across calls) */
	float *ws_feat = nullptr, *ws_y = nullptr;
	double *ws_lith = nullptr, *ws_dP = nullptr, *ws_dv = nullptr;
	double *cb_dirs = nullptr, *cb_ys = nullptr;
	int cb_k = -1;
	long reorth = 0;         /* inner Gram-Schmidt second passes (diagnostic) */
	float *st12 = nullptr;   /* strain rates and stresses, 12 level-0 planes */
};

static int choose_bottom(i3g_ctx *c);

/* ---- event timing of named kernels ------------------------------------ */

static hipEvent_t ev_get(i3g_ctx *c)
{
	if (c->evnext >= c->evpool.size()) {
		hipEvent_t e;
		hipEventCreate(&e);
		c->evpool.push_back(e);
	}
	return c->evpool[c->evnext++];
}

/* bytes of sweep operands per z-slab of the automatic slab choice */
#define SLAB_BYTES 300e6

/* Event pair around the launches of one scope, kept under `name` when
 * timing is on (a null name times nothing) */
struct Timed {
	i3g_ctx *c;
	const char *name;
	double bytes;
	hipEvent_t a, b;
	Timed(i3g_ctx *c_, const char *n, double by) : c(c_), name(n), bytes(by)
	{
		if (c->timing && name) {
			a = ev_get(c);
			b = ev_get(c);
			hipEventRecord(a, c->s);
		}
	}
	~Timed()
	{
		if (c->timing && name) {
			hipEventRecord(b, c->s);
			c->pending.push_back({name, {a, b}});
			c->stats[name].bytes += bytes;
		}
	}
};

static void flush_timing(i3g_ctx *c)
{
	if (!c->timing) return;
	hipStreamSynchronize(c->s);
	for (auto &p : c->pending) {
		float ms = 0;
		hipEventElapsedTime(&ms, p.second.first, p.second.second);
		KStat &k = c->stats[p.first];
		k.ms += ms;
		k.calls += 1;
	}
	c->pending.clear();
	c->evnext = 0;
}

/* ------------------------------------------------------------------ */
/* Setup                                                              */
/* ------------------------------------------------------------------ */

static inline dim3 lin_grid(long n) { return dim3((unsigned)((n + 255) / 256)); }

static inline dim3 plane_grid(int nx, int ny, int nz)
{
	long pl = (long)nx * ny;
	return dim3((unsigned)((pl + 255) / 256), (unsigned)nz);
}
/* one thread per node and unknown kind: z of the grid is the kind */
static inline dim3 comp_grid(int nx, int ny, int nz)
{
	long pl = (long)nx * ny;
	return dim3((unsigned)((pl + 255) / 256), (unsigned)nz, 4);
}
static inline dim3 half_grid(int nx, int ny, int nz)
{
	long pl = (long)nx * ((ny + 1) / 2);
	return dim3((unsigned)((pl + 255) / 256), (unsigned)nz);
}

static int exists_uk(int c, long m1, long m2, long m3, long nx, long ny, long nz)
{
	switch (c) {
	case 0: return m1 && m2 && m3;
	case 1: return m2 < ny - 1 && m3 < nz - 1;
	case 2: return m1 < nx - 1 && m3 < nz - 1;
	default: return m1 < nx - 1 && m2 < ny - 1;
	}
}

/* Coarse base index and 1D weights of a fine index for the transfer kinds
 * of vprestrict/vpprolong (move3mg.c:2776-2789, 2806-2823, 3201-3214) and
 * of ronurestrict (kind 3, move3mg.c:2670-2712). */
static long xbase(long m, long nf, long nc, const double *gf, const double *gc, int kind)
{
	long r = (m - 2) / 2 + 2;
	if (m < 3) r = m;
	if (m > nf - 4) r = nc - (nf - m);
	if (kind == 0 || kind == 3) {
		if (r > nc - 2) r = nc - 2;
	} else if (kind == 1) {
		if (r > nc - 3) r = nc - 3;
		if ((gf[m] + gf[m + 1]) / 2.0 < (gc[r] + gc[r + 1]) / 2.0 && r > 0) r -= 1;
	} else {
		if (r > nc - 2) r = nc - 2;
		if ((gf[m] + gf[m - 1]) / 2.0 < (gc[r] + gc[r - 1]) / 2.0 && r > 1) r -= 1;
	}
	return r;
}

static double xw1(long m, long base, long cc, const double *gf, const double *gc, int kind)
{
	double d, dd;
	if (kind == 3) {
		d = gc[base + 1] - gc[base];
		dd = (gc[cc] - gf[m]) / d;
		return 1.0 - fabs(dd);
	}
	if (kind == 0) {
		d = gc[base + 1] - gc[base];
		dd = (gc[cc] - gf[m]) / d;
	} else if (kind == 1) {
		d = (gc[base + 2] - gc[base]) / 2.0;
		dd = (gc[cc] + gc[cc + 1] - gf[m] - gf[m + 1]) / 2.0 / d;
	} else {
		d = (gc[base + 1] - gc[base - 1]) / 2.0;
		dd = (gc[cc] + gc[cc - 1] - gf[m] - gf[m - 1]) / 2.0 / d;
	}
	if (cc == base) dd = -dd;
	return 1.0 - dd;
}

/* Position of the point of a transfer kind on a line set: node kinds at
 * the line, mid kind between lines m and m+1, centre kind between m-1 and m */
static double kind_pos(const double *g, long m, int kind)
{
	if (kind == 1) return (g[m] + g[m + 1]) / 2.0;
	if (kind == 2) return (g[m - 1] + g[m]) / 2.0;
	return g[m];
}

/* Geometric alternative to xbase: the coarse pair that brackets the fine
 * point (largest admissible base whose coarse point is not above it).
 * Equal to xbase wherever the coarse lines are every second fine line;
 * differs next to the copied outer lines when (n-5) is odd, where xbase
 * pairs a fine line with a coarse pair that does not contain it. */
static long xbase_geo(long m, long nc, const double *gf, const double *gc, int kind)
{
	long lo = kind == 2 ? 1 : 0, hi = kind == 1 ? nc - 3 : nc - 2;
	double p = kind_pos(gf, m, kind);
	long b = lo;
	for (long c = lo; c <= hi; c++)
		if (kind_pos(gc, c, kind) <= p) b = c;
	return b;
}

static int build_axis(AxisMapDev &A, long nf, long nc, const double *gf, const double *gc, int geo)
{
	int err = 0;
	for (int kind = 0; kind < 4; kind++) {
		/* fine indices for which this kind is defined: node kinds all; mid
		 * kind needs m < nf-1; centre kind needs m >= 1 */
		std::vector<int> base(nf, 0);
		std::vector<double> wlo(nf, 0), whi(nf, 0);
		std::vector<std::vector<std::pair<int, double>>> inv(nc);
		for (long m = 0; m < nf; m++) {
			if (kind == 1 && m >= nf - 1) continue;
			if (kind == 2 && m < 1) continue;
			long b = geo ? xbase_geo(m, nc, gf, gc, kind) : xbase(m, nf, nc, gf, gc, kind);
			if (b < 0 || b + 1 >= nc || (kind == 1 && b + 2 >= nc) || (kind == 2 && b < 1))
				return I3G_EINVAL;
			base[m] = (int)b;
			wlo[m] = xw1(m, b, b, gf, gc, kind);
			whi[m] = xw1(m, b, b + 1, gf, gc, kind);
			if (geo) {
				/* interpolation inside the pair, the nearest coarse
				 * point beyond the outermost one */
				wlo[m] = wlo[m] < 0 ? 0 : wlo[m] > 1 ? 1 : wlo[m];
				whi[m] = whi[m] < 0 ? 0 : whi[m] > 1 ? 1 : whi[m];
			}
			inv[b].push_back({(int)m, wlo[m]});
			inv[b + 1].push_back({(int)m, whi[m]});
		}
		std::vector<int> st(nc + 1, 0), idx;
		std::vector<double> w;
		for (long c = 0; c < nc; c++) {
			st[c] = (int)idx.size();
			if (inv[c].size() > XFER_MAXL) return I3G_EINVAL; /* k_restrict's register list */
			double wsum = 0, wabs = 0;
			for (auto &p : inv[c]) {
				idx.push_back(p.first);
				w.push_back(p.second);
				wsum += p.second;
				wabs += fabs(p.second);
			}
			/* the node-viscosity restriction divides by the product of the
			 * three 1-D weight sums; a vanishing sum (the outer-line rule
			 * of an even-sized level can give weights -1, 0, 1, 0) would
			 * make the coarse viscosity non-finite */
			if (kind == 3 && fabs(wsum) <= 1e-9 * wabs) return I3G_EINVAL;
		}
		st[nc] = (int)idx.size();
		std::vector<float> wlo32(wlo.begin(), wlo.end()), whi32(whi.begin(), whi.end()), w32(w.begin(), w.end());
		A.base[kind] = dalloc<int>(nf, &err);
		A.wlo64[kind] = dalloc<double>(nf, &err);
		A.whi64[kind] = dalloc<double>(nf, &err);
		A.wlo32[kind] = dalloc<float>(nf, &err);
		A.whi32[kind] = dalloc<float>(nf, &err);
		A.istart[kind] = dalloc<int>(nc + 1, &err);
		A.iidx[kind] = dalloc<int>(idx.size(), &err);
		A.iw64[kind] = dalloc<double>(w.size(), &err);
		A.iw32[kind] = dalloc<float>(w.size(), &err);
		if (err) return err;
		if (upload(A.base[kind], base.data(), nf) || upload(A.wlo64[kind], wlo.data(), nf) ||
		    upload(A.whi64[kind], whi.data(), nf) || upload(A.wlo32[kind], wlo32.data(), nf) ||
		    upload(A.whi32[kind], whi32.data(), nf) || upload(A.istart[kind], st.data(), nc + 1) ||
		    upload(A.iidx[kind], idx.data(), idx.size()) || upload(A.iw64[kind], w.data(), w.size()) ||
		    upload(A.iw32[kind], w32.data(), w.size()))
			return I3G_EDEVICE;
	}
	return I3G_OK;
}

static int build_level(i3g_ctx *c, int l, const i3g_setup *s)
{
	Level &L = c->L[l];
	const i3g_hier &h = c->h;
	int err = 0;
	L.nx = (int)h.mgx[l]; L.ny = (int)h.mgy[l]; L.nz = (int)h.mgz[l];
	L.N = h.mgn[l];
	long N = L.N, nx = L.nx, ny = L.ny, nz = L.nz;
	for (int a = 0; a < 3; a++) L.hg[a].resize(a == 0 ? nx : a == 1 ? ny : nz);
	i3g_level_lines(&h, l, s->gx, s->gy, s->gz, L.hg[0].data(), L.hg[1].data(), L.hg[2].data());
	for (int a = 0; a < 3; a++) {
		long n = (long)L.hg[a].size();
		const double *g = L.hg[a].data();
		std::vector<double> ih(n, 0), ic(n, 0);
		for (long i = 0; i + 1 < n; i++) ih[i] = 1.0 / (g[i + 1] - g[i]);
		for (long i = 1; i + 1 < n; i++) ic[i] = 1.0 / (g[i + 1] - g[i - 1]);
		std::vector<float> ih32(ih.begin(), ih.end()), ic32(ic.begin(), ic.end());
		L.g[a] = dalloc<double>(n, &err);
		L.ih64[a] = dalloc<double>(n, &err);
		L.ic64[a] = dalloc<double>(n, &err);
		L.ih32[a] = dalloc<float>(n, &err);
		L.ic32[a] = dalloc<float>(n, &err);
		if (err) return err;
		if (upload(L.g[a], g, n) || upload(L.ih64[a], ih.data(), n) || upload(L.ic64[a], ic.data(), n) ||
		    upload(L.ih32[a], ih32.data(), n) || upload(L.ic32[a], ic32.data(), n))
			return I3G_EDEVICE;
	}
	/* flags and boundary records */
	std::vector<uint8_t> fl(N, 0);
	std::vector<int> bcid(4 * N, -1);
	struct R { int u, p; double c, k; int depth; };
	std::vector<R> recs;
	long mgp = h.mgp[l];
	for (long m3 = 0; m3 < nz; m3++)
		for (long m1 = 0; m1 < nx; m1++)
			for (long m2 = 0; m2 < ny; m2++) {
				long n = m3 * nx * ny + m1 * ny + m2;
				for (int cc = 0; cc < 4; cc++) {
					if (!exists_uk(cc, m1, m2, m3, nx, ny, nz)) continue;
					fl[n] |= FL_EX(cc);
					int rec = s->bondm[(mgp + n) * 4 + cc];
					if (!rec) continue;
					if (rec < 0 || rec >= s->nrec) return I3G_EBC;
					fl[n] |= FL_BC(cc);
					R r;
					r.u = (int)(cc * N + n);
					r.c = s->bondv[2 * rec];
					r.k = s->bondv[2 * rec + 1];
					r.p = -1;
					r.depth = 0;
					if (s->bondn[rec]) {
						long g = (long)s->bondn[rec] - 1;
						long gn = g / 4 - mgp;
						int gc = (int)(g % 4);
						if (gn < 0 || gn >= N) return I3G_EBC;
						r.p = (int)(gc * N + gn);
					}
					bcid[r.u] = (int)recs.size();
					recs.push_back(r);
				}
			}
	/* the interior stencil rows must not reach outside the grid: every
	 * unknown on the rims where a stencil is undefined has to be a BC row */
	for (long m3 = 0; m3 < nz; m3++)
		for (long m1 = 0; m1 < nx; m1++)
			for (long m2 = 0; m2 < ny; m2++) {
				long n = m3 * nx * ny + m1 * ny + m2;
				for (int cc = 1; cc < 4; cc++) {
					if (!(fl[n] & FL_EX(cc)) || (fl[n] & FL_BC(cc))) continue;
					bool ok = m1 >= 1 && m2 >= 1 && m3 >= 1 && m1 < nx - 1 && m2 < ny - 1 && m3 < nz - 1;
					if (cc == 1) ok = ok && m2 < ny - 2 && m3 < nz - 2;
					if (cc == 2) ok = ok && m1 < nx - 2 && m3 < nz - 2;
					if (cc == 3) ok = ok && m1 < nx - 2 && m2 < ny - 2;
					if (!ok) return I3G_EBC;
				}
			}
	/* dependency depth of each BC row (partner chains) */
	for (size_t q = 0; q < recs.size(); q++) {
		int d = 1, u = (int)q;
		while (recs[u].p >= 0 && bcid[recs[u].p] >= 0) {
			u = bcid[recs[u].p];
			if (++d > 64) return I3G_EBC; /* cycle */
		}
		recs[q].depth = d;
	}
	std::vector<int> order(recs.size());
	for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return recs[a].depth < recs[b].depth; });
	std::vector<int> bu(recs.size()), bp(recs.size());
	std::vector<double> bcv(recs.size()), bkv(recs.size());
	int maxd = 0;
	for (size_t q = 0; q < order.size(); q++) {
		const R &r = recs[order[q]];
		bu[q] = r.u; bp[q] = r.p; bcv[q] = r.c; bkv[q] = r.k;
		bcid[r.u] = (int)q;
		maxd = std::max(maxd, r.depth);
	}
	L.depth_start.assign(maxd + 2, 0);
	for (size_t q = 0; q < order.size(); q++) L.depth_start[recs[order[q]].depth + 1] = (int)q + 1;
	for (int d = 1; d <= maxd + 1; d++) L.depth_start[d] = std::max(L.depth_start[d], L.depth_start[d - 1]);
	L.nbc = (int)recs.size();
	std::vector<float> bcv32(bcv.begin(), bcv.end()), bkv32(bkv.begin(), bkv.end());
	L.flags = dalloc<uint8_t>(N, &err);
	L.bcid = dalloc<int>(4 * N, &err);
	L.bc_u = dalloc<int>(L.nbc, &err);
	L.bc_p = dalloc<int>(L.nbc, &err);
	L.bc_c64 = dalloc<double>(L.nbc, &err);
	L.bc_k64 = dalloc<double>(L.nbc, &err);
	L.bc_c32 = dalloc<float>(L.nbc, &err);
	L.bc_k32 = dalloc<float>(L.nbc, &err);
	L.eta64 = dalloc<double>(4 * N, &err);
	L.eta32 = dalloc<float>(4 * N, &err);
	L.wP64 = dalloc<double>(N, &err);
	L.wP32 = dalloc<float>(N, &err);
	L.nu = dalloc<float>(N, &err);
	L.x64 = dalloc<double>(4 * N, &err);
	L.b64 = dalloc<double>(4 * N, &err);
	L.r64 = dalloc<double>(4 * N, &err);
	L.x32 = dalloc<float>(4 * N, &err);
	L.b32 = dalloc<float>(4 * N, &err);
	L.r32 = dalloc<float>(4 * N, &err);
	L.x64b = dalloc<double>(4 * N, &err);
	L.x32b = dalloc<float>(4 * N, &err);
	if (err) return err;
	{
		/* balanced tiles of about ftile[] nodes (x, y) and slabs of about
		 * ftile[2] planes */
		FusedGeom &G = L.fg;
		G.ntx = (int)((nx + c->ftile[0] - 1) / c->ftile[0]);
		G.nty = (int)((ny + c->ftile[1] - 1) / c->ftile[1]);
		G.nzs = (int)((nz + c->ftile[2] - 1) / c->ftile[2]);
		G.txm = (int)((nx + G.ntx - 1) / G.ntx);
		G.tym = (int)((ny + G.nty - 1) / G.nty);
		G.remap = 0;
	}
	if (upload(L.flags, fl.data(), N) || upload(L.bcid, bcid.data(), 4 * N) ||
	    upload(L.bc_u, bu.data(), L.nbc) || upload(L.bc_p, bp.data(), L.nbc) ||
	    upload(L.bc_c64, bcv.data(), L.nbc) || upload(L.bc_k64, bkv.data(), L.nbc) ||
	    upload(L.bc_c32, bcv32.data(), L.nbc) || upload(L.bc_k32, bkv32.data(), L.nbc))
		return I3G_EDEVICE;
	HCHK(hipMemset(L.x64, 0, 4 * N * sizeof(double)));
	HCHK(hipMemset(L.x32, 0, 4 * N * sizeof(float)));
	HCHK(hipMemset(L.x64b, 0, 4 * N * sizeof(double)));
	HCHK(hipMemset(L.x32b, 0, 4 * N * sizeof(float)));
	return I3G_OK;
}

template <typename T> static LevelView<T> view(const Level &L);
template <> LevelView<double> view<double>(const Level &L)
{
	LevelView<double> v;
	v.nx = L.nx; v.ny = L.ny; v.nz = L.nz; v.N = L.N;
	v.ihx = L.ih64[0]; v.ihy = L.ih64[1]; v.ihz = L.ih64[2];
	v.icx = L.ic64[0]; v.icy = L.ic64[1]; v.icz = L.ic64[2];
	v.eC = L.eta64; v.eXY = L.eta64 + L.N; v.eXZ = L.eta64 + 2 * L.N; v.eYZ = L.eta64 + 3 * L.N;
	v.flags = L.flags; v.wP = L.wP64;
	return v;
}
template <> LevelView<float> view<float>(const Level &L)
{
	LevelView<float> v;
	v.nx = L.nx; v.ny = L.ny; v.nz = L.nz; v.N = L.N;
	v.ihx = L.ih32[0]; v.ihy = L.ih32[1]; v.ihz = L.ih32[2];
	v.icx = L.ic32[0]; v.icy = L.ic32[1]; v.icz = L.ic32[2];
	v.eC = L.eta32; v.eXY = L.eta32 + L.N; v.eXZ = L.eta32 + 2 * L.N; v.eYZ = L.eta32 + 3 * L.N;
	v.flags = L.flags; v.wP = L.wP32;
	return v;
}
template <typename T> static BcRec<T> bcview(const Level &L);
template <> BcRec<double> bcview<double>(const Level &L) { return {L.bc_u, L.bc_p, L.bc_c64, L.bc_k64}; }
template <> BcRec<float> bcview<float>(const Level &L) { return {L.bc_u, L.bc_p, L.bc_c32, L.bc_k32}; }
template <typename T> static XferView<T> xview(const Xfer &X);
template <> XferView<double> xview<double>(const Xfer &X)
{
	XferView<double> v;
	for (int a = 0; a < 3; a++)
		for (int q = 0; q < 3; q++) {
			v.ax[a].base[q] = X.ax[a].base[q];
			v.ax[a].wlo[q] = X.ax[a].wlo64[q];
			v.ax[a].whi[q] = X.ax[a].whi64[q];
			v.ax[a].istart[q] = X.ax[a].istart[q];
			v.ax[a].iidx[q] = X.ax[a].iidx[q];
			v.ax[a].iw[q] = X.ax[a].iw64[q];
		}
	return v;
}
template <> XferView<float> xview<float>(const Xfer &X)
{
	XferView<float> v;
	for (int a = 0; a < 3; a++)
		for (int q = 0; q < 3; q++) {
			v.ax[a].base[q] = X.ax[a].base[q];
			v.ax[a].wlo[q] = X.ax[a].wlo32[q];
			v.ax[a].whi[q] = X.ax[a].whi32[q];
			v.ax[a].istart[q] = X.ax[a].istart[q];
			v.ax[a].iidx[q] = X.ax[a].iidx[q];
			v.ax[a].iw[q] = X.ax[a].iw32[q];
		}
	return v;
}
template <typename T> static T *lx(Level &L);
template <> double *lx<double>(Level &L) { return L.x64; }
template <> float *lx<float>(Level &L) { return L.x32; }
template <typename T> static T *lxb(Level &L);
template <> double *lxb<double>(Level &L) { return L.x64b; }
template <> float *lxb<float>(Level &L) { return L.x32b; }
template <typename T> static T *lb(Level &L);
template <> double *lb<double>(Level &L) { return L.b64; }
template <> float *lb<float>(Level &L) { return L.b32; }
template <typename T> static T *lr(Level &L);
template <> double *lr<double>(Level &L) { return L.r64; }
template <> float *lr<float>(Level &L) { return L.r32; }

/* ---- scratch of the entry point that allocates nothing ------------------- */

/* Device memory for i3g_stokes_residual_noalloc: the per-block partial sums
 * of the residual norms and their final reduction.  Every GPU the library
 * runs on holds its own copy until the process ends.  In host code the
 * array's name is not a device address: the address of each device's copy
 * is looked up once per device (scratch_init, from i3g_create with that
 * device current) and only that pointer is handed to kernels. */
#define I3G_SCRATCH_BYTES ((size_t)4 << 20)
#define I3G_SCRATCH_MAXDEV 64
static __device__ __align__(256) unsigned char scratch[I3G_SCRATCH_BYTES];
static std::mutex scratch_mu;
static void *scratch_dev[I3G_SCRATCH_MAXDEV];

static int scratch_init(int dev)
{
	if (dev < 0 || dev >= I3G_SCRATCH_MAXDEV) return I3G_EDEVICE;
	std::lock_guard<std::mutex> g(scratch_mu);
	if (scratch_dev[dev]) return I3G_OK;
	void *ptr = nullptr;
	if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return I3G_EDEVICE;
	scratch_dev[dev] = ptr;
	return I3G_OK;
}

extern "C" int i3g_create(const i3g_setup *s, i3g_ctx **out)
{
	if (!s || !out || !s->gx || !s->gy || !s->gz || !s->bondm || !s->bondv || !s->bondn)
		return I3G_EINVAL;
	*out = nullptr;
	i3g_ctx *c = new i3g_ctx();
	int rc = i3g_hierarchy(s->nx, s->ny, s->nz, s->multinum, &c->h);
	if (rc) { delete c; return rc; }
	if (s->nx * s->ny * s->nz > 400000000L) { delete c; return I3G_ECAP; }
	c->nlev = c->h.nlev;
	{
		const char *e = getenv("I3G_XCD_REMAP");
		if (e && e[0] >= '0' && e[0] <= '2') c->xcd_remap = e[0] - '0';
		e = getenv("I3G_SMOOTH_ILP");
		if (e && e[0] == '0') c->smooth_ilp = 0;
		e = getenv("I3G_SLAB");
		if (e && e[0]) c->slab = atoi(e) > 0 ? atoi(e) : 0;
		e = getenv("I3G_SLAB_ALL");
		if (e && e[0] == '1') c->slab_all = 1;
		e = getenv("I3G_COARSE_VISC");
		if (e && e[0] >= '0' && e[0] <= '2') c->coarse_visc = e[0] - '0';
		e = getenv("I3G_COARSE_VISC_MAXLEV");
		if (e) c->coarse_visc_maxlev = atoi(e);
		e = getenv("I3G_XFER_GEO");
		if (e && e[0] == '1') c->xfer_geo = 1;
		e = getenv("I3G_FUSED_BOTTOM");
		if (e && e[0] == '0') c->bottom = -1;
		e = getenv("I3G_BC_CHAIN");
		if (e && e[0] == '0') c->bc_chain = 0;
		e = getenv("I3G_FUSED_SWEEP");
		if (e) c->fused = e[0] == '1';
		e = getenv("I3G_SWEEP7");
		if (e) c->sweep7 = e[0] == '1';
		e = getenv("I3G_COARSE_DIRECT");
		if (e) c->coarse_direct = e[0] == '1';
		e = getenv("I3G_SCGS");
		if (e) c->scgs = e[0] == '1';
		e = getenv("I3G_SCGS_LEVELS"); /* "last" or "first,last" */
		if (e) {
			int a = 0, b = 0;
			if (sscanf(e, "%d,%d", &a, &b) == 2) { c->scgs_minlev = a; c->scgs_maxlev = b; }
			else c->scgs_maxlev = atoi(e);
		}
		e = getenv("I3G_SCGS_MIX");
		if (e) c->scgs_mix = atoi(e);
		e = getenv("I3G_SCGS_PHASES");
		if (e) c->scgs_phases = atoi(e);
		e = getenv("I3G_SCGS_GUARD");
		if (e) c->scgs_guard = atoi(e);
		e = getenv("I3G_SCGS_GUARD_RATIO");
		if (e && atof(e) > 0) c->scgs_guard_ratio = atof(e);
		e = getenv("I3G_SCGS_W");
		if (e) {
			double a = 0, b = 0;
			if (sscanf(e, "%lf,%lf", &a, &b) == 2 && a > 0 && b > 0) { c->scgs_wv = a; c->scgs_wp = b; }
		}
		e = getenv("I3G_COARSE_DIRECT_TOL");
		if (e && atof(e) > 0) c->cd_tol = atof(e);
		e = getenv("I3G_COARSEST64");
		if (e) c->coarsest64 = e[0] != '0';
		e = getenv("I3G_SWEEP3");
		if (e) c->sweep3 = e[0] == '1';
		e = getenv("I3G_SWEEP3_MIN_GROUPS");
		if (e) c->sweep3_min_groups = atoi(e);
		e = getenv("I3G_SWEEP3_NZS");
		if (e) c->sweep3_nzs = atoi(e);
		e = getenv("I3G_SWEEP7_MIN_TILES");
		if (e) c->sweep7_min_tiles = atoi(e);
		e = getenv("I3G_FUSED_MINZ");
		if (e) c->fused_minz = atoi(e);
		e = getenv("I3G_FUSED_TILE");
		if (e) {
			int a = 0, b = 0, d = 0;
			if (sscanf(e, "%d,%d,%d", &a, &b, &d) == 3 && a > 0 && b > 0 && d > 0) {
				c->ftile[0] = a; c->ftile[1] = b; c->ftile[2] = d;
			}
		}
	}
	if (hipGetDevice(&c->dev) != hipSuccess || scratch_init(c->dev) != I3G_OK) { delete c; return I3G_EDEVICE; }
	if (hipStreamCreate(&c->s) != hipSuccess) { delete c; return I3G_EDEVICE; }
	c->L.resize(c->nlev);
	for (int l = 0; l < c->nlev; l++) {
		rc = build_level(c, l, s);
		if (rc) { i3g_destroy(c); return rc; }
	}
	c->X.resize(c->nlev > 1 ? c->nlev - 1 : 0);
	for (int l = 0; l + 1 < c->nlev; l++) {
		Level &F = c->L[l], &Cl = c->L[l + 1];
		for (int a = 0; a < 3; a++) {
			long nf = (long)F.hg[a].size(), nc = (long)Cl.hg[a].size();
			rc = build_axis(c->X[l].ax[a], nf, nc, F.hg[a].data(), Cl.hg[a].data(), c->xfer_geo);
			if (rc) { i3g_destroy(c); return rc; }
		}
	}
	if (c->bottom == 0) c->bottom = choose_bottom(c);
	else c->bottom = 0;
	int err = 0;
	if (c->scgs)
		for (int l = c->scgs_minlev; l < c->nlev && l <= c->scgs_maxlev; l++)
			c->L[l].scgs = dalloc<double>(7 * c->L[l].N, &err);
	long N0 = c->L[0].N;
	c->rhs0 = dalloc<double>(4 * N0, &err);
	c->sol = dalloc<double>(4 * N0, &err);
	c->ro = dalloc<float>(N0, &err);
	c->nxx = dalloc<float>(N0, &err);
	c->nxy = dalloc<float>(N0, &err);
	c->nxz = dalloc<float>(N0, &err);
	c->nyz = dalloc<float>(N0, &err);
	c->npart = 8192;
	c->partials = dalloc<double>(c->npart * 2, &err);
	c->dscal = dalloc<double>(64, &err);
	c->e64 = dalloc<double>(4 * N0, &err);
	if (err) { i3g_destroy(c); return err; }
	if (hipHostMalloc((void **)&c->hscal, 4096 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
		i3g_destroy(c);
		return I3G_ENOMEM;
	}
	hipMemset(c->sol, 0, 4 * N0 * sizeof(double));
	*out = c;
	return I3G_OK;
}

static void free_krylov(i3g_ctx *c)
{
	for (void *p : c->P) hipFree(p);
	for (void *p : c->Q) hipFree(p);
	c->P.clear(); c->Q.clear();
	void *ps[] = {c->kr, c->ke, c->dP, c->dQ, c->dh};
	for (void *p : ps)
		if (p) hipFree(p);
	c->kr = c->ke = nullptr;
	c->dP = c->dQ = nullptr;
	c->dh = nullptr;
	c->kr_m = 0;
	c->kr_k = 0;
	for (void *q : c->iV) hipFree(q);
	for (void *q : c->iZ) hipFree(q);
	c->iV.clear(); c->iZ.clear();
	if (c->diV) hipFree(c->diV);
	if (c->diZ) hipFree(c->diZ);
	c->diV = c->diZ = nullptr;
	c->in_m = 0;
}

extern "C" int i3g_destroy(i3g_ctx *c)
{
	if (!c) return I3G_OK;
	hipDeviceSynchronize();
	free_krylov(c);
	for (int q = 0; q < 2; q++) {
		if (c->cg_exec[q]) hipGraphExecDestroy(c->cg_exec[q]);
		if (c->cg_graph[q]) hipGraphDestroy(c->cg_graph[q]);
	}
	for (auto &L : c->L) {
		void *ps[] = {L.g[0], L.g[1], L.g[2], L.ih64[0], L.ih64[1], L.ih64[2], L.ic64[0], L.ic64[1],
			      L.ic64[2], L.ih32[0], L.ih32[1], L.ih32[2], L.ic32[0], L.ic32[1], L.ic32[2],
			      L.eta64, L.eta32, L.wP64, L.wP32, L.flags, L.bcid, L.bc_u, L.bc_p, L.bc_c64,
			      L.bc_k64, L.bc_c32, L.bc_k32, L.nu, L.x64, L.b64, L.r64, L.x32, L.b32, L.r32,
			      L.x64b, L.x32b, L.scgs};
		for (void *p : ps)
			if (p) hipFree(p);
	}
	for (auto &X : c->X)
		for (int a = 0; a < 3; a++)
			for (int q = 0; q < 4; q++) {
				void *ps[] = {X.ax[a].base[q], X.ax[a].wlo64[q], X.ax[a].whi64[q], X.ax[a].iw64[q],
					      X.ax[a].wlo32[q], X.ax[a].whi32[q], X.ax[a].iw32[q], X.ax[a].istart[q],
					      X.ax[a].iidx[q]};
				for (void *p : ps)
					if (p) hipFree(p);
			}
	void *ps[] = {c->rhs0, c->sol, c->ro, c->nxx, c->nxy, c->nxz, c->nyz, c->partials, c->dscal, c->e64, c->st12,
		      c->ws_feat, c->ws_y, c->ws_lith,
		      c->ws_dP, c->ws_dv, c->cb_dirs, c->cb_ys, c->cd_u, c->cd_vt, c->cd_s, c->cd_y};
	for (void *p : ps)
		if (p) hipFree(p);
	if (c->hscal) hipHostFree(c->hscal);
	for (auto e : c->evpool) hipEventDestroy(e);
	if (c->s) hipStreamDestroy(c->s);
	delete c;
	return I3G_OK;
}

/* ---- the coarsest level's direct FP64 solve ---------------------------- */

__global__ void k_unit(double *x, long j, long jprev)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		if (jprev >= 0) x[jprev] = 0;
		x[j] = 1;
	}
}
/* column j of the column-major operator from the residual of e_j (r = -A e_j) */
__global__ void k_dense_col(double *A, const double *r, long n, long j)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) A[j * n + i] = -r[i];
}
/* unknowns that do not exist on the staggered grid: identity rows */
__global__ void k_dense_ident(double *A, const uint8_t *flags, long N)
{
	long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (u >= 4 * N) return;
	int c = (int)(u / N);
	if (!(flags[u - c * N] & FL_EX(c))) A[u * 4 * N + u] = 1.0;
}
/* out[k] = scale[k] * sum_i M[k n + i] v[i] (M column-major: column k of M
 * against v); one wavefront per k, fixed summation order */
__global__ void __launch_bounds__(256) k_gemvt(const double *__restrict__ M, const double *__restrict__ v,
					       const double *__restrict__ scale, double *__restrict__ out, int n)
{
	int k = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (k >= n) return;
	const double *col = M + (long)k * n;
	double acc = 0;
	for (int i = lane; i < n; i += 64) acc += col[i] * v[i];
	for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
	if (lane == 0) out[k] = scale ? scale[k] * acc : acc;
}

/* Assemble the coarsest level's operator (BC rows x_u - k x_partner,
 * identity rows for unknowns that do not exist), factor it by SVD
 * (rocSOLVER, FP64) and keep U, V^T and the reciprocal singular values
 * above cd_tol times the largest: the V-cycle's coarsest solve is then
 * x = V S^+ U^T b, the minimum-norm least-squares solution, which leaves
 * the operator's null and near-null modes (constant pressure, pressure
 * cells the coarse viscosities decouple) out of the correction. */
static int coarse_direct_setup(i3g_ctx *c)
{
	Level &L = c->L[c->nlev - 1];
	long n = 4 * L.N;
	if (n > 16384) return I3G_ECAP;
	double t0 = (double)clock() / CLOCKS_PER_SEC;
	int err = 0;
	if (c->cd_n != (int)n) {
		for (void *p : {(void *)c->cd_u, (void *)c->cd_vt, (void *)c->cd_s, (void *)c->cd_y})
			if (p) hipFree(p);
		c->cd_u = dalloc<double>(n * n, &err);
		c->cd_vt = dalloc<double>(n * n, &err);
		c->cd_s = dalloc<double>(n, &err);
		c->cd_y = dalloc<double>(n, &err);
		if (err) return err;
		c->cd_n = (int)n;
	}
	double *A = dalloc<double>(n * n, &err);
	double *E = dalloc<double>(n, &err);
	int *info = dalloc<int>(1, &err);
	if (err) return err;
	HCHK(hipMemsetAsync(L.x64, 0, n * sizeof(double), c->s));
	for (long j = 0; j < n; j++) {
		hipLaunchKernelGGL(k_unit, dim3(1), dim3(64), 0, c->s, L.x64, j, j - 1);
		hipLaunchKernelGGL((k_residual<double, double, 0>), plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s,
				   view<double>(L), (const double *)L.x64, (const double *)nullptr, L.r64, L.bcid,
				   bcview<double>(L), (double *)nullptr);
		hipLaunchKernelGGL(k_dense_col, lin_grid(n), dim3(256), 0, c->s, A, (const double *)L.r64, n, j);
	}
	hipLaunchKernelGGL(k_dense_ident, lin_grid(n), dim3(256), 0, c->s, A, (const uint8_t *)L.flags, L.N);
	HCHK(hipGetLastError());
	rocblas_handle h = nullptr;
	int rc = I3G_OK;
	if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, c->s) != rocblas_status_success ||
	    rocsolver_dgesvd(h, rocblas_svect_all, rocblas_svect_all, (rocblas_int)n, (rocblas_int)n, A, (rocblas_int)n,
			     c->cd_s, c->cd_u, (rocblas_int)n, c->cd_vt, (rocblas_int)n, E, rocblas_outofplace,
			     info) != rocblas_status_success)
		rc = I3G_EDEVICE;
	int hinfo = 0;
	std::vector<double> sv(n);
	if (!rc) {
		HCHK(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->s));
		HCHK(hipMemcpyAsync(sv.data(), c->cd_s, n * sizeof(double), hipMemcpyDeviceToHost, c->s));
	}
	HCHK(hipStreamSynchronize(c->s));
	if (h) rocblas_destroy_handle(h);
	hipFree(A);
	hipFree(E);
	hipFree(info);
	if (rc) return rc;
	if (hinfo != 0) return I3G_ENOCONV;
	int rank = 0;
	const double smax = sv[0];
	for (long k = 0; k < n; k++) {
		bool keep = sv[k] > c->cd_tol * smax;
		sv[k] = keep ? 1.0 / sv[k] : 0.0;
		rank += keep;
	}
	if (upload(c->cd_s, sv.data(), n)) return I3G_EDEVICE;
	c->cd_rank = rank;
	c->cd_setup_s = (double)clock() / CLOCKS_PER_SEC - t0;
	fprintf(stderr, "i3g coarse direct: %ld unknowns, rank %d (tol %.1e), setup %.2f s CPU\n", n, rank, c->cd_tol,
		c->cd_setup_s);
	if (getenv("I3G_COARSE_DIRECT_DIAG")) {
		fprintf(stderr, "  largest singular value %.3e, smallest kept %.3e\n", smax,
			rank > 0 ? 1.0 / sv[rank - 1] : 0.0);
	}
	return I3G_OK;
}

extern "C" int i3g_set_coefficients(i3g_ctx *c, const float *nu, const float *nxx, const float *nxy,
				    const float *nxz, const float *nyz, const float *ro, const double *phys)
{
	if (!c || !nu || !nxx || !nxy || !nxz || !nyz || !ro || !phys) return I3G_EINVAL;
	for (int q = 0; q < 6; q++) c->phys[q] = phys[q];
	int viscmod = (int)phys[5];
	if (viscmod < 0 || viscmod > 2) return I3G_EINVAL;
	long N0 = c->L[0].N;
	HCHK(hipMemcpyAsync(c->L[0].nu, nu, N0 * sizeof(float), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->nxx, nxx, N0 * sizeof(float), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->nxy, nxy, N0 * sizeof(float), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->nxz, nxz, N0 * sizeof(float), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->nyz, nyz, N0 * sizeof(float), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->ro, ro, N0 * sizeof(float), hipMemcpyDefault, c->s));
	/* viscmod selects the arithmetic, geometric or harmonic averages of
	 * the level-0 stress-point viscosities and of ronurestrict, as in the
	 * reference; I3G_COARSE_VISC overrides the coarse levels' choice */
	int cm0 = c->coarse_visc >= 0 ? c->coarse_visc : viscmod;
	auto cmode = [&](int l) { return l <= c->coarse_visc_maxlev ? cm0 : viscmod; };
	for (int l = 1; l < c->nlev; l++) {
		Level &F = c->L[l - 1], &Cl = c->L[l];
		const Xfer &X = c->X[l - 1];
		int cm = cmode(l);
		hipLaunchKernelGGL(k_ronurestrict, plane_grid(Cl.nx, Cl.ny, Cl.nz), dim3(256), 0, c->s,
				   F.nx, F.ny, Cl.nx, Cl.ny, Cl.nz, F.nu, Cl.nu, X.ax[0].istart[3],
				   X.ax[0].iidx[3], X.ax[0].iw64[3], X.ax[1].istart[3], X.ax[1].iidx[3],
				   X.ax[1].iw64[3], X.ax[2].istart[3], X.ax[2].iidx[3], X.ax[2].iw64[3], cm);
	}
	for (int l = 0; l < c->nlev; l++) {
		Level &L = c->L[l];
		hipLaunchKernelGGL(k_eta, plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s, L.nx, L.ny, L.nz,
				   l == 0 ? 1 : 0, L.nu, c->nxx, c->nxy, c->nxz, c->nyz, l == 0 ? viscmod : cmode(l), phys[3],
				   phys[4], L.eta64, L.eta32, L.wP64, L.wP32, L.g[0], L.g[1], L.g[2]);
	}
	Level &L0 = c->L[0];
	hipLaunchKernelGGL(k_rhs, plane_grid(L0.nx, L0.ny, L0.nz), dim3(256), 0, c->s, L0.nx, L0.ny, L0.nz,
			   c->ro, phys[0], phys[1], phys[2], L0.flags, L0.bcid, L0.bc_c64, c->rhs0);
	HCHK(hipGetLastError());
	HCHK(hipStreamSynchronize(c->s));
	c->cd_rank = 0;
	if (c->coarse_direct && c->nlev > 1) {
		int rc = coarse_direct_setup(c);
		if (rc) return rc;
	}
	c->coef_ready = 1;
	return I3G_OK;
}

/* ------------------------------------------------------------------ */
/* Multigrid                                                          */
/* ------------------------------------------------------------------ */

template <typename T> static void bc_enforce(i3g_ctx *c, Level &L, T *x, int homog)
{
	BcRec<T> bc = bcview<T>(L);
	int nd = (int)L.depth_start.size() - 2;
	if (c->bc_chain && nd <= BC_MAXCHAIN) {
		int a = L.depth_start[1], b = L.depth_start[nd + 1];
		if (b > a)
			hipLaunchKernelGGL(k_bc_chain<T>, dim3((b - a + 255) / 256), dim3(256), 0, c->s, x, bc,
					   (const int *)L.bcid, a, b, homog);
		return;
	}
	for (size_t d = 1; d + 1 < L.depth_start.size(); d++) {
		int a = L.depth_start[d], b = L.depth_start[d + 1];
		if (b > a)
			hipLaunchKernelGGL(k_bc<T>, dim3((b - a + 255) / 256), dim3(256), 0, c->s, x, bc, a, b, homog);
	}
}

template <typename T>
static void sweep(i3g_ctx *c, int l, T *x, const T *b, const i3g_stokes_params *p)
{
	Level &L = c->L[l];
	LevelView<T> v = view<T>(L);
	dim3 hg = half_grid(L.nx, L.ny, L.nz);
	int chunks = (int)hg.x, remap = c->xcd_remap && (l == 0 || (c->xcd_remap == 2 && L.nz >= 64));
	const bool f64 = sizeof(T) == 8;
	/* The seven phases (vx, vy, vz in two colours each, then P) in z-slabs:
	 * in slab s (planes [b_s, b_s+1)) phase q relaxes planes
	 * [b_s - q, b_s+1 - q), the last slab up to the top.  A row reaches
	 * one plane up and down, so phase q at plane k finds phase q-1 done at
	 * k+1 (and below), phase q+1 not yet begun at k-1, and the later
	 * components and P untouched at k and k+1: every node is relaxed from
	 * the same values as in the phase-by-phase sweep, and the result is
	 * the same to the bit.  What it buys: the slab's planes, read from HBM
	 * by the first phase, are still in the last-level cache for the other
	 * six. */
	const int nq = 7;
	int slab = c->slab;
	if (slab < 0) {
		/* the operands of a sweep (four unknown and four rhs planes, four
		 * viscosity planes, the flag byte) in slabs that stay in the 256 MB
		 * last-level cache with the lagging phases' planes: three slabs at
		 * 257^3 in FP32, one (no slabs) at 129^3 */
		double bytes = (double)L.N * (12 * sizeof(T) + 1);
		int ns = (int)ceil(bytes / SLAB_BYTES);
		slab = ns > 1 ? (L.nz + ns - 1) / ns : 0;
	}
	if (!(slab > 0 && L.nz > slab + nq && (l == 0 || c->slab_all))) slab = L.nz;
	double by = slab < L.nz ? 0 : (double)L.N * (sizeof(T) * 2 + 1 + sizeof(T) * 4 + sizeof(T) * 3) / 2;
	{
	Timed tsw(c, slab < L.nz ? (l == 0 ? (f64 ? "sweep_slab_f64_L0" : "sweep_slab_f32_L0") : "sweep_slab_coarse")
				 : nullptr,
		  0);
	for (int b0 = 0; b0 < L.nz; b0 += slab) {
		int b1 = b0 + slab >= L.nz ? L.nz : b0 + slab;
		for (int q = 0; q < nq; q++) {
			int k0 = b0 - q < 0 ? 0 : b0 - q;
			int k1 = b1 == L.nz ? L.nz : b1 - q;
			int np = k1 - k0;
			if (np <= 0) continue;
			if (q == 6) {
				Timed tm(c, slab < L.nz ? nullptr : (l == 0 ? (f64 ? "smooth_p_f64_L0" : "smooth_p_f32_L0")
									      : "smooth_p_coarse"),
					 0);
				dim3 pg = plane_grid(L.nx, L.ny, np);
				hipLaunchKernelGGL(k_smooth_p<T>, pg, dim3(256), 0, c->s, v, x, b, (T)p->p0koef, k0);
				continue;
			}
			int comp = 1 + q / 2, col = q & 1;
			dim3 g = remap ? dim3((unsigned)((((long)chunks * np + 7) / 8) * 8)) : dim3((unsigned)chunks, (unsigned)np);
			Timed tm(c, slab < L.nz ? nullptr : (l == 0 ? (f64 ? "smooth_vel_f64_L0" : "smooth_vel_f32_L0")
								      : "smooth_vel_coarse"),
				 by);
			if (!c->smooth_ilp)
				hipLaunchKernelGGL(k_smooth_vel<T>, g, dim3(256), 0, c->s, v, x, b, comp, col, (T)p->v0koef,
						   remap, chunks, k0, np);
			else if (comp == 1)
				hipLaunchKernelGGL((k_smooth_vel_ilp<T, 1>), g, dim3(256), 0, c->s, v, x, b, col,
						   (T)p->v0koef, remap, chunks, k0, np);
			else if (comp == 2)
				hipLaunchKernelGGL((k_smooth_vel_ilp<T, 2>), g, dim3(256), 0, c->s, v, x, b, col,
						   (T)p->v0koef, remap, chunks, k0, np);
			else
				hipLaunchKernelGGL((k_smooth_vel_ilp<T, 3>), g, dim3(256), 0, c->s, v, x, b, col,
						   (T)p->v0koef, remap, chunks, k0, np);
		}
	}
	}
	{
		Timed tm(c, l == 0 ? (f64 ? "bc_f64_L0" : "bc_f32_L0") : "bc_coarse", 0);
		bc_enforce<T>(c, L, x, 1);
	}
}

/* One sweep xin -> xout in three launches (k_sweep_comp): vx, vy, vz with
 * P; then the BC rows of xout.  Equals sweep() on xin to the bit. */
template <typename T> static size_t fused_lds(const Level &L)
{
	return (size_t)4 * (L.fg.txm + 4) * (L.fg.tym + 4) * sizeof(T);
}

template <typename T>
static void fused_sweep(i3g_ctx *c, int l, const T *xin, T *xout, const T *b, const i3g_stokes_params *p)
{
	Level &L = c->L[l];
	LevelView<T> v = view<T>(L);
	long N = L.N;
	FusedGeom G = L.fg;
	G.remap = c->xcd_remap && (l == 0 || (c->xcd_remap == 2 && L.nz >= 64));
	int total = G.ntx * G.nty * G.nzs;
	dim3 grid = G.remap ? dim3((unsigned)(((total + 7) / 8) * 8)) : dim3((unsigned)total);
	size_t lds = fused_lds<T>(L);
	if (lds > 64 * 1024) {
		/* opt in to the larger LDS allocation once per kernel */
		static size_t granted[2] = {0, 0};
		size_t &gr = granted[sizeof(T) == 8];
		if (lds > gr) {
			hipFuncSetAttribute((const void *)k_sweep_comp<T, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
			hipFuncSetAttribute((const void *)k_sweep_comp<T, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
			hipFuncSetAttribute((const void *)k_sweep_comp<T, 3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
			gr = lds;
		}
	}
	T v0 = (T)p->v0koef, p0 = (T)p->p0koef;
	/* compulsory bytes of the component launches: own component read and
	 * written, the other three unknown planes, three viscosity planes, the
	 * rhs and the flag byte (the vz launch adds P written and its rhs) */
	double w = sizeof(T), by = (double)N * (2 * w + 3 * w + 3 * w + w + 1);
	const char *n1 = l == 0 ? (sizeof(T) == 8 ? "sweep_vx_f64_L0" : "sweep_vx_f32_L0") : "sweep_coarse";
	const char *n2 = l == 0 ? (sizeof(T) == 8 ? "sweep_vy_f64_L0" : "sweep_vy_f32_L0") : "sweep_coarse";
	const char *n3 = l == 0 ? (sizeof(T) == 8 ? "sweep_vzp_f64_L0" : "sweep_vzp_f32_L0") : "sweep_coarse";
	{
		Timed tm(c, n1, by);
		hipLaunchKernelGGL((k_sweep_comp<T, 1>), grid, dim3(256), lds, c->s, v, xin, xin + N, xin + 2 * N,
				   xin + 3 * N, b, xout + N, (T *)nullptr, G, v0, p0);
	}
	{
		Timed tm(c, n2, by);
		hipLaunchKernelGGL((k_sweep_comp<T, 2>), grid, dim3(256), lds, c->s, v, xin, (const T *)(xout + N),
				   xin + 2 * N, xin + 3 * N, b, xout + 2 * N, (T *)nullptr, G, v0, p0);
	}
	{
		Timed tm(c, n3, by + (double)N * 2 * w);
		hipLaunchKernelGGL((k_sweep_comp<T, 3>), grid, dim3(256), lds, c->s, v, xin, (const T *)(xout + N),
				   (const T *)(xout + 2 * N), xin + 3 * N, b, xout + 3 * N, xout, G, v0, p0);
	}
	bc_enforce<T>(c, L, xout, 1);
}

/* k_sweep7 tiling of a level: tiles of at most SW7_T lines in x and y,
 * enough of them to fill the device; false when the level keeps the
 * phase launches */
static bool sw7_geom(i3g_ctx *c, int l, Sw7Geom *G)
{
	const Level &L = c->L[l];
	if (!c->sweep7 || L.nz > SW7_MAXZ || L.nx < 3 || L.ny < 3) return false;
	int ntx = (L.nx + SW7_T - 1) / SW7_T, nty = (L.ny + SW7_T - 1) / SW7_T;
	if (ntx * nty < c->sweep7_min_tiles) return false;
	size_t lds = sw7_lds_bytes(L.nz);
	if (lds > c->sweep7_lds) {
		int dev = 0, lim = 0;
		if (hipGetDevice(&dev) != hipSuccess ||
		    hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
		    (size_t)lim < lds ||
		    hipFuncSetAttribute((const void *)k_sweep7, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
			    hipSuccess) {
			hipGetLastError();
			c->sweep7 = 0;
			return false;
		}
		c->sweep7_lds = lds;
	}
	G->ntx = ntx;
	G->nty = nty;
	G->remap = c->xcd_remap ? 1 : 0;
	return true;
}

/* One FP32 sweep xin -> xout in one launch (k_sweep7), then the BC rows of
 * xout.  Equals sweep() on xin to the bit. */
static void sweep7(i3g_ctx *c, int l, const float *xin, float *xout, const float *b, const i3g_stokes_params *p,
		   const Sw7Geom &G)
{
	Level &L = c->L[l];
	int total = G.ntx * G.nty;
	dim3 grid = G.remap ? dim3((unsigned)(((total + 7) / 8) * 8)) : dim3((unsigned)total);
	{
		Timed tm(c, l == 0 ? "sweep7_f32_L0" : "sweep7_coarse", 0);
		hipLaunchKernelGGL(k_sweep7, grid, dim3(SW7_THREADS), sw7_lds_bytes(L.nz), c->s, view<float>(L), xin,
				   xout, b, G, (float)p->v0koef, (float)p->p0koef);
	}
	{
		Timed tm(c, l == 0 ? "bc_f32_L0" : "bc_coarse", 0);
		bc_enforce<float>(c, L, xout, 1);
	}
}

/* k_sweep3 tiling of a level: tiles of at most SW3_TX x SW3_TY lines and
 * z-slabs of at least 8 planes, as many as fill two work-groups per
 * compute unit; false when the level keeps the phase launches */
static bool sw3_geom(i3g_ctx *c, int l, Sw3Geom *G)
{
	const Level &L = c->L[l];
	if (!c->sweep3 || L.nz > 4096 || L.nx < 3 || L.ny < 3 || L.nz < 3) return false;
	int ntx = (L.nx + SW3_TX - 1) / SW3_TX, nty = (L.ny + SW3_TY - 1) / SW3_TY;
	int nzs = c->sweep3_nzs > 0 ? c->sweep3_nzs : 512 / (ntx * nty);
	if (nzs > L.nz / 8) nzs = L.nz / 8;
	if (nzs < 1) nzs = 1;
	if (ntx * nty * nzs < c->sweep3_min_groups) return false;
	size_t lds = sw3_lds_bytes(3, L.nz);
	if (sw3_lds_bytes(1, L.nz) > lds) lds = sw3_lds_bytes(1, L.nz);
	if (sw3_lds_bytes(2, L.nz) > lds) lds = sw3_lds_bytes(2, L.nz);
	if (lds > c->sweep3_lds) {
		int dev = 0, lim = 0;
		if (hipGetDevice(&dev) != hipSuccess ||
		    hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
		    (size_t)lim < lds ||
		    hipFuncSetAttribute((const void *)k_sweep3<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
			    hipSuccess ||
		    hipFuncSetAttribute((const void *)k_sweep3<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
			    hipSuccess ||
		    hipFuncSetAttribute((const void *)k_sweep3<3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
			    hipSuccess) {
			hipGetLastError();
			c->sweep3 = 0;
			return false;
		}
		c->sweep3_lds = lds;
	}
	G->ntx = ntx;
	G->nty = nty;
	G->nzs = nzs;
	G->remap = c->xcd_remap ? 1 : 0;
	return true;
}

/* One FP32 sweep xin -> xout in three launches (k_sweep3<1..3>), then the
 * BC rows of xout.  Equals sweep() on xin to the bit. */
static void sweep3(i3g_ctx *c, int l, const float *xin, float *xout, const float *b, const i3g_stokes_params *p,
		   const Sw3Geom &G)
{
	Level &L = c->L[l];
	int total = G.ntx * G.nty * G.nzs;
	dim3 grid = G.remap ? dim3((unsigned)(((total + 7) / 8) * 8)) : dim3((unsigned)total);
	LevelView<float> v = view<float>(L);
	float v0 = (float)p->v0koef, p0 = (float)p->p0koef;
	{
		Timed tm(c, l == 0 ? "sweep3_f32_L0" : "sweep3_coarse", 0);
		hipLaunchKernelGGL(k_sweep3<1>, grid, dim3(SW3_THREADS), sw3_lds_bytes(1, L.nz), c->s, v, xin,
				   (const float *)xout, xout, b, G, v0, p0);
		hipLaunchKernelGGL(k_sweep3<2>, grid, dim3(SW3_THREADS), sw3_lds_bytes(2, L.nz), c->s, v, xin,
				   (const float *)xout, xout, b, G, v0, p0);
		hipLaunchKernelGGL(k_sweep3<3>, grid, dim3(SW3_THREADS), sw3_lds_bytes(3, L.nz), c->s, v, xin,
				   (const float *)xout, xout, b, G, v0, p0);
	}
	{
		Timed tm(c, l == 0 ? "bc_f32_L0" : "bc_coarse", 0);
		bc_enforce<float>(c, L, xout, 1);
	}
}

/* One sweep of the coupled cell smoother: both colours, then the BC rows */
template <typename T> static void scgs_sweep(i3g_ctx *c, int l, T *x, const T *b)
{
	Level &L = c->L[l];
	LevelView<T> v = view<T>(L);
	/* blocks over (chunk, plane), remapped so that each XCD takes a
	 * contiguous run of planes (block_work) on level 0 */
	int remap = c->xcd_remap && l == 0;
	int ch_c = (int)(((long)L.nx * ((L.ny + 1) / 2) + 255) / 256), ch_a = (int)(((long)L.nx * L.ny + 255) / 256);
	auto grid_of = [&](int chunks) {
		return remap ? dim3((unsigned)((((long)chunks * L.nz + 7) / 8) * 8)) : dim3((unsigned)chunks, (unsigned)L.nz);
	};
	dim3 gc = grid_of(ch_c), ga = grid_of(ch_a);
	{
		Timed tm(c, l == 0 ? "scgs_L0" : "scgs_coarse", 0);
		for (int col = 0; col < 2; col++) {
			hipLaunchKernelGGL(k_scgs_cell<T>, gc, dim3(256), 0, c->s, v, (const T *)x, b, (T *)L.scgs, col, remap,
					   ch_c);
			hipLaunchKernelGGL(k_scgs_apply<T>, ga, dim3(256), 0, c->s, v, x, (const T *)L.scgs, col,
					   (T)c->scgs_wv, (T)c->scgs_wp, remap, ch_a);
		}
	}
	bc_enforce<T>(c, L, x, 1);
}

/* ns sweeps on the level's unknowns: pairs of out-of-place sweeps through
 * the second vector (so the result lands back in x), an odd one in place */
template <typename T> static void smooth_phases(i3g_ctx *c, int l, int ns, const i3g_stokes_params *p);

template <typename T> static void smooth(i3g_ctx *c, int l, int ns, const i3g_stokes_params *p)
{
	Level &L = c->L[l];
	if (!L.scgs || c->scgs_off) {
		smooth_phases<T>(c, l, ns, p);
		return;
	}
	T *x = lx<T>(L), *b = lb<T>(L);
	if (c->scgs_mix > 0) {
		smooth_phases<T>(c, l, c->scgs_phases > 0 && c->scgs_phases < ns ? c->scgs_phases : ns, p);
		for (int s = 0; s < c->scgs_mix; s++) scgs_sweep<T>(c, l, x, b);
		return;
	}
	for (int s = 0; s < ns; s++) scgs_sweep<T>(c, l, x, b);
}

template <typename T> static void smooth_phases(i3g_ctx *c, int l, int ns, const i3g_stokes_params *p)
{
	Level &L = c->L[l];
	T *x = lx<T>(L), *xb = lxb<T>(L), *b = lb<T>(L);
	int s = 0;
	if constexpr (sizeof(T) == 4) {
		Sw7Geom G;
		Sw3Geom G3;
		if (sw7_geom(c, l, &G))
			for (; s + 2 <= ns; s += 2) {
				sweep7(c, l, x, xb, b, p, G);
				sweep7(c, l, xb, x, b, p, G);
			}
		else if (sw3_geom(c, l, &G3))
			for (; s + 2 <= ns; s += 2) {
				sweep3(c, l, x, xb, b, p, G3);
				sweep3(c, l, xb, x, b, p, G3);
			}
	}
	if (c->fused && L.nz >= c->fused_minz)
		for (; s + 2 <= ns; s += 2) {
			fused_sweep<T>(c, l, x, xb, b, p);
			fused_sweep<T>(c, l, xb, x, b, p);
		}
	for (; s < ns; s++) sweep<T>(c, l, x, b, p);
}

template <typename T> static void residual_level(i3g_ctx *c, int l, const T *x, const T *b, T *r)
{
	Level &L = c->L[l];
	Timed tm(c, l == 0 ? (sizeof(T) == 8 ? "residual_f64_L0" : "residual_f32_L0") : "residual_coarse", 0);
	hipLaunchKernelGGL((k_residual<T, T, 0>), plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s,
			   view<T>(L), x, b, r, L.bcid, bcview<T>(L), (double *)nullptr);
}

/* V-cycle on level l: x_l (in/out), b_l given (level buffers). */
template <typename T> static void vcycle(i3g_ctx *c, int l, const i3g_stokes_params *p);

static size_t coarsest_lds(const Level &L, size_t word)
{
	return (size_t)L.N * (8 * word + 1);
}

/* LDS budget of the coarsest-level kernel: the device's per-work-group
 * limit, with the opt-in above 64 KiB set for both precisions; 0 = off */
static int choose_bottom(i3g_ctx *c)
{
	const Level &L = c->L[c->nlev - 1];
	if (c->nlev < 2 || L.depth_start.size() > COARSEST_MAXD + 2) return 0;
	int dev = 0, lim = 0;
	if (hipGetDevice(&dev) != hipSuccess ||
	    hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
		return 0;
	size_t need = coarsest_lds(L, sizeof(double));
	if (need > (size_t)lim) need = coarsest_lds(L, sizeof(float));
	if (need > (size_t)lim) return 0;
	if (need > 64 * 1024) {
		if (hipFuncSetAttribute((const void *)k_coarsest<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
					(int)need) != hipSuccess ||
		    hipFuncSetAttribute((const void *)k_coarsest<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
					(int)need) != hipSuccess) {
			hipGetLastError();
			need = 64 * 1024;
		}
	}
	c->coarsest_lds_max = need;
	return 1;
}

/* the coarsest level's sweeps in one launch when its data fits the LDS
 * budget in precision T; false: use the launched sweeps */
template <typename T> static bool coarsest_launch(i3g_ctx *c, const i3g_stokes_params *p, int sweeps)
{
	Level &L = c->L[c->nlev - 1];
	size_t lds = coarsest_lds(L, sizeof(T));
	if (c->bottom <= 0 || lds > c->coarsest_lds_max) return false;
	CoarsestArgs<T> A;
	memset(&A, 0, sizeof(A));
	A.v = view<T>(L);
	A.x = lx<T>(L);
	A.b = lb<T>(L);
	A.eta = A.v.eC;
	A.bc = bcview<T>(L);
	A.nd = (int)L.depth_start.size() - 2;
	for (size_t d = 0; d < L.depth_start.size(); d++) A.ds[d] = L.depth_start[d];
	A.sweeps = sweeps;
	A.v0koef = (T)p->v0koef;
	A.p0koef = (T)p->p0koef;
	hipLaunchKernelGGL(k_coarsest<T>, dim3(1), dim3(COARSEST_THREADS), lds, c->s, A);
	return true;
}

/* The coarse levels (>= 1) of a V-cycle do many small launches; they are
 * captured once into a HIP graph (fixed level buffers, no memory-set
 * nodes) and replayed.  Re-captured when the smoothing parameters change. */
template <typename T> static void coarse_cycle_graph(i3g_ctx *c, const i3g_stokes_params *p)
{
	int q = sizeof(T) == 8 ? 1 : 0;
	std::vector<double> key = {p->p0koef, p->p1koef, p->v0koef, p->v1koef, (double)p->nu_coarse, (double)c->scgs_off};
	for (int l = 1; l < c->nlev; l++) { key.push_back(p->nu_pre[l]); key.push_back(p->nu_post[l]); }
	if (!c->cg_exec[q] || c->cg_key[q] != key) {
		if (c->cg_exec[q]) { hipGraphExecDestroy(c->cg_exec[q]); c->cg_exec[q] = nullptr; }
		if (c->cg_graph[q]) { hipGraphDestroy(c->cg_graph[q]); c->cg_graph[q] = nullptr; }
		bool tm = c->timing;
		c->timing = false;
		bool ok = hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal) == hipSuccess;
		if (ok) {
			vcycle<T>(c, 1, p);
			ok = hipStreamEndCapture(c->s, &c->cg_graph[q]) == hipSuccess;
		}
		c->timing = tm;
		if (ok) ok = hipGraphInstantiate(&c->cg_exec[q], c->cg_graph[q], nullptr, nullptr, 0) == hipSuccess;
		if (!ok) {
			/* fall back to direct launches for this context */
			c->cg_exec[q] = nullptr;
			c->use_graph = 0;
			hipGetLastError();
			vcycle<T>(c, 1, p);
			return;
		}
		c->cg_key[q] = key;
	}
	Timed tm(c, "coarse_levels_graph", 0);
	hipGraphLaunch(c->cg_exec[q], c->s);
}

template <typename T> static void vcycle(i3g_ctx *c, int l, const i3g_stokes_params *p)
{
	Level &L = c->L[l];
	T *x = lx<T>(L), *b = lb<T>(L);
	if (l == c->nlev - 1) {
		int ns = p->nu_coarse > 0 ? p->nu_coarse : 64;
		if (c->cd_rank > 0) {
			/* direct FP64 solve: x = V S^+ U^T b */
			long n4 = 4 * L.N;
			const double *bb = (const double *)b;
			if (sizeof(T) == 4) {
				hipLaunchKernelGGL((k_copy_scale<double, T>), lin_grid(n4), dim3(256), 0, c->s, L.b64,
						   (const T *)b, 1.0, n4);
				bb = L.b64;
			}
			dim3 g((unsigned)((n4 + 3) / 4));
			hipLaunchKernelGGL(k_gemvt, g, dim3(256), 0, c->s, (const double *)c->cd_u, bb, (const double *)c->cd_s,
					   c->cd_y, (int)n4);
			if (sizeof(T) == 4) {
				hipLaunchKernelGGL(k_gemvt, g, dim3(256), 0, c->s, (const double *)c->cd_vt,
						   (const double *)c->cd_y, (const double *)nullptr, L.x64, (int)n4);
				hipLaunchKernelGGL((k_copy_scale<T, double>), lin_grid(n4), dim3(256), 0, c->s, x,
						   (const double *)L.x64, 1.0, n4);
			} else {
				hipLaunchKernelGGL(k_gemvt, g, dim3(256), 0, c->s, (const double *)c->cd_vt,
						   (const double *)c->cd_y, (const double *)nullptr, (double *)x, (int)n4);
			}
			return;
		}
		if (sizeof(T) == 4 && c->coarsest64) {
			/* the mixed V-cycle's coarsest solve in FP64: its right-hand
			 * side and start widened, the sweeps in double, the result
			 * rounded back */
			long n4 = 4 * L.N;
			Timed tm(c, "coarsest_f64", 0);
			hipLaunchKernelGGL((k_copy_scale<double, T>), lin_grid(n4), dim3(256), 0, c->s, L.b64, (const T *)b, 1.0,
					   n4);
			hipLaunchKernelGGL((k_copy_scale<double, T>), lin_grid(n4), dim3(256), 0, c->s, L.x64, (const T *)x, 1.0,
					   n4);
			if (!coarsest_launch<double>(c, p, ns)) smooth<double>(c, l, ns, p);
			hipLaunchKernelGGL((k_copy_scale<T, double>), lin_grid(n4), dim3(256), 0, c->s, x, (const double *)L.x64,
					   1.0, n4);
			return;
		}
		if (coarsest_launch<T>(c, p, ns)) return;
		smooth<T>(c, l, ns, p);
		return;
	}
	smooth<T>(c, l, p->nu_pre[l], p);
	residual_level<T>(c, l, x, b, lr<T>(L));
	Level &Cl = c->L[l + 1];
	{
		Timed tm(c, "restrict", 0);
		hipLaunchKernelGGL(k_restrict<T>, comp_grid(Cl.nx, Cl.ny, Cl.nz), dim3(256), 0, c->s, view<T>(L),
				   view<T>(Cl), xview<T>(c->X[l]), (const T *)lr<T>(L), lb<T>(Cl), lx<T>(Cl));
	}
	if (l == 0 && c->use_graph) coarse_cycle_graph<T>(c, p);
	else vcycle<T>(c, l + 1, p);
	{
		Timed tm(c, "prolong", 0);
		hipLaunchKernelGGL(k_prolong<T>, plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s, view<T>(L),
				   view<T>(Cl), xview<T>(c->X[l]), (const T *)lx<T>(Cl), x, (T)p->v1koef,
				   (T)p->p1koef);
	}
	smooth<T>(c, l, p->nu_post[l], p);
}

/* ---- reductions -------------------------------------------------------- */

static int dot_blocks(long n)
{
	long b = (n + 4095) / 4096;
	if (b > 2048) b = 2048;
	if (b < 1) b = 1;
	return (int)b;
}

template <typename A, typename B> static double dot(i3g_ctx *c, const A *a, const B *b, long n)
{
	int nb = dot_blocks(n);
	hipLaunchKernelGGL((k_dot_part<A, B>), dim3(nb), dim3(256), 0, c->s, a, b, n, c->partials);
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->s, c->partials, (long)nb, 1, c->dscal);
	hipMemcpyAsync(c->hscal, c->dscal, sizeof(double), hipMemcpyDeviceToHost, c->s);
	hipStreamSynchronize(c->s);
	return c->hscal[0];
}

/* FP64 residual of the level-0 system into r (may be L0.r64) and the
 * weighted metric: out[0] = ||W r|| over non-BC rows, out[1] = ||W b||. */
static int resid64(i3g_ctx *c, const double *x, double *r, double *out)
{
	Level &L = c->L[0];
	dim3 g = plane_grid(L.nx, L.ny, L.nz);
	long nb = (long)g.x * g.y;
	if (nb > c->npart) {
		hipFree(c->partials);
		int err = 0;
		c->npart = nb;
		c->partials = dalloc<double>(2 * nb, &err);
		if (err) return err;
	}
	{
		Timed tm(c, "residual_f64_L0", (double)L.N * (32 + 32 + 32 + 8 + 1 + 32));
		hipLaunchKernelGGL((k_residual<double, double, 0>), g, dim3(256), 0, c->s, view<double>(L), x,
				   (const double *)c->rhs0, r, L.bcid, bcview<double>(L), c->partials);
	}
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->s, c->partials, nb, 2, c->dscal);
	HCHK(hipMemcpyAsync(c->hscal, c->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
	HCHK(hipStreamSynchronize(c->s));
	out[0] = sqrt(c->hscal[0]);
	out[1] = sqrt(c->hscal[1]);
	return I3G_OK;
}

extern "C" int i3g_stokes_residual(i3g_ctx *c, const double *P, const double *vx, const double *vy,
				   const double *vz, double *rP, double *rvx, double *rvy, double *rvz,
				   double *norms)
{
	if (!c || !c->coef_ready || !P || !vx || !vy || !vz) return I3G_EINVAL;
	Level &L = c->L[0];
	long N = L.N;
	const double *src[4] = {P, vx, vy, vz};
	for (int q = 0; q < 4; q++)
		HCHK(hipMemcpyAsync(c->e64 + q * N, src[q], N * sizeof(double), hipMemcpyHostToDevice, c->s));
	double nrm[2];
	int rc = resid64(c, c->e64, L.r64, nrm);
	if (rc) return rc;
	double *dst[4] = {rP, rvx, rvy, rvz};
	for (int q = 0; q < 4; q++)
		if (dst[q]) HCHK(hipMemcpyAsync(dst[q], L.r64 + q * N, N * sizeof(double), hipMemcpyDeviceToHost, c->s));
	HCHK(hipStreamSynchronize(c->s));
	if (norms) { norms[0] = nrm[0]; norms[1] = nrm[1]; }
	return I3G_OK;
}

__global__ void k_sqrt2(const double *in, double *out)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		out[0] = sqrt(in[0]);
		out[1] = sqrt(in[1]);
	}
}

/* i3g_stokes_residual on device buffers with nothing allocated: the block
 * partial sums and their reduction live in the static scratch (at most
 * I3G_SCRATCH_BYTES: 2 doubles per block of 256 nodes, plus 2), the
 * residual goes to d_r or, when d_r is NULL, to the level's own work
 * vector.  Same kernels and block partition as i3g_stokes_residual, so the
 * results are the same to the bit.  Enqueued on `stream` (NULL: the
 * context's stream), not synchronised.  Calls that overlap on one GPU share
 * its scratch and race. */
extern "C" int i3g_stokes_residual_noalloc(i3g_ctx *c, const double *d_x, double *d_r, double *d_norms, void *stream)
{
	if (!c || !c->coef_ready || !d_x || !d_norms) return I3G_EINVAL;
	if (c->dev < 0 || c->dev >= I3G_SCRATCH_MAXDEV || !scratch_dev[c->dev]) return I3G_EDEVICE;
	Level &L = c->L[0];
	dim3 g = plane_grid(L.nx, L.ny, L.nz);
	long nb = (long)g.x * g.y;
	if ((size_t)(2 * nb + 2) * sizeof(double) > I3G_SCRATCH_BYTES) return I3G_ECAP;
	double *part = (double *)scratch_dev[c->dev];
	hipStream_t st = stream ? (hipStream_t)stream : c->s;
	double *r = d_r ? d_r : L.r64;
	hipLaunchKernelGGL((k_residual<double, double, 0>), g, dim3(256), 0, st, view<double>(L), d_x,
			   (const double *)c->rhs0, r, L.bcid, bcview<double>(L), part);
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double *)part, nb, 2, part + 2 * nb);
	hipLaunchKernelGGL(k_sqrt2, dim3(1), dim3(64), 0, st, (const double *)(part + 2 * nb), d_norms);
	HCHK(hipGetLastError());
	return I3G_OK;
}

/* ------------------------------------------------------------------ */
/* FGMRES                                                             */
/* ------------------------------------------------------------------ */


template <typename T> static int ensure_krylov(i3g_ctx *c, int m)
{
	if (c->kr_m == m && c->kr_prec == (int)sizeof(T)) return I3G_OK;
	free_krylov(c);
	long n = 4 * c->L[0].N;
	int err = 0;
	/* keep the direction storage within 80% of the free device memory */
	size_t fr = 0, tot = 0;
	if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
		long cap = (long)(0.8 * (double)fr / (2.0 * (double)n * sizeof(T))) - 2;
		if (cap < 4) return I3G_ENOMEM;
		if (m > cap) m = (int)cap;
	}
	for (int q = 0; q < m; q++) c->P.push_back(dalloc<T>(n, &err));
	for (int q = 0; q < m; q++) c->Q.push_back(dalloc<T>(n, &err));
	c->kr = dalloc<double>(n, &err);
	c->ke = dalloc<double>(n, &err);
	c->dP = dalloc<void *>(m, &err);
	c->dQ = dalloc<void *>(m, &err);
	c->dh = dalloc<double>(3 * m + 16, &err);
	if (err) { free_krylov(c); return err; }
	if (upload(c->dP, c->P.data(), m) || upload(c->dQ, c->Q.data(), m)) { free_krylov(c); return I3G_EDEVICE; }
	if (c->npart < 1024L * m) {
		hipFree(c->partials);
		c->npart = 1024L * m;
		c->partials = dalloc<double>(2 * c->npart, &err);
		if (err) { free_krylov(c); return err; }
	}
	c->kr_m = m;
	c->kr_prec = (int)sizeof(T);
	c->kr_k = 0;
	return I3G_OK;
}

/* scale the P plane by wP (fwd) or 1/wP (inverse); other planes copied.
 * Boundary-condition rows are set to 0: they are satisfied exactly by
 * every iterate and correction, so they are kept out of the Krylov space. */
template <typename TO, typename TI>
__global__ void k_wscale(TO *__restrict__ y, const TI *__restrict__ x, const double *__restrict__ wP,
			 const uint8_t *__restrict__ flags, long N, int inv, double a)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= 4 * N) return;
	long n = t % N;
	int c = (int)(t / N);
	double v = (double)x[t] * a;
	if (c == 0) {
		double w = wP[n];
		v = inv ? (w != 0.0 ? v / w : 0.0) : v * w;
	}
	if (flags[n] & FL_BC(c)) v = 0;
	y[t] = (TO)v;
}

/* Block Gram-Schmidt helpers over j stored vectors V[0..j): partial dot
 * products of q with every V[i] (groups of 8 vectors per pass over the
 * block's chunk), then y += sign * sum_i h[i] V[i]. */
template <typename T, typename TQ>
__global__ void __launch_bounds__(256) k_multidot(const T *const *V, int j, const TQ *__restrict__ q, long n,
						  double *partials)
{
	__shared__ double s[8][256];
	long chunk = (n + gridDim.x - 1) / gridDim.x;
	long lo = (long)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
	for (int g0 = 0; g0 < j; g0 += 8) {
		int g1 = g0 + 8 < j ? g0 + 8 : j;
		double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		for (long t = lo + threadIdx.x; t < hi; t += blockDim.x) {
			double qv = (double)q[t];
#pragma unroll
			for (int i = 0; i < 8; i++)
				if (g0 + i < g1) acc[i] += (double)V[g0 + i][t] * qv;
		}
#pragma unroll
		for (int i = 0; i < 8; i++) s[i][threadIdx.x] = acc[i];
		__syncthreads();
		for (int w = 128; w > 0; w >>= 1) {
			if (threadIdx.x < w)
#pragma unroll
				for (int i = 0; i < 8; i++) s[i][threadIdx.x] += s[i][threadIdx.x + w];
			__syncthreads();
		}
		if (threadIdx.x < g1 - g0) partials[(long)blockIdx.x * j + g0 + threadIdx.x] = s[threadIdx.x][0];
		__syncthreads();
	}
}

template <typename TY, typename T>
__global__ void k_multiaxpy(TY *__restrict__ y, const T *const *V, const double *__restrict__ h, int j,
			    long n, double sign)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double acc = 0;
	for (int i = 0; i < j; i++) acc += h[i] * (double)V[i][t];
	y[t] = (TY)((double)y[t] + sign * acc);
}

/* y *= 1/sqrt(s[0]) (s = squared norm on the device) */
template <typename T> __global__ void k_scale_rsqrt(T *__restrict__ y, const double *s, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double f = s[0] > 0 ? 1.0 / sqrt(s[0]) : 0.0;
	if (t < n) y[t] = (T)((double)y[t] * f);
}

/* y += sign * a[0] * x */
template <typename TY, typename T>
__global__ void k_axpy_dev(TY *__restrict__ y, const T *__restrict__ x, const double *a, double sign, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) y[t] = (TY)((double)y[t] + sign * a[0] * (double)x[t]);
}

/* device dot products into out[0..j) (no host synchronisation) */
template <typename T, typename TQ>
static void multidot_dev(i3g_ctx *c, const T *const *V, int j, const TQ *q, long n, double *out)
{
	int nb = dot_blocks(n);
	if (nb > 1024) nb = 1024;
	hipLaunchKernelGGL((k_multidot<T, TQ>), dim3(nb), dim3(256), 0, c->s, V, j, q, n, c->partials);
	hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->s, c->partials, (long)nb, j, out);
}

/* GCR on (W A) e = r in the W-scaled space with the V-cycle as (right,
 * flexible) preconditioner.  Raw directions p_j = M W^-1 r_j are kept with
 * orthonormal q_j, W A P = Q R (R upper triangular, on the host), so only
 * q is orthogonalised; the correction is e = P R^-1 alpha.  Directions are
 * kept across the outer refinement steps of one solve (recycling: a new
 * residual is first projected on span(Q)).  The operator product W A p is
 * evaluated in FP64 even when p and the V-cycle are FP32.  r: scaled
 * residual (in/out); e: unscaled correction (accumulated).  Returns the
 * V-cycles applied; *red = final ||r|| / initial ||r||. */
struct GcrHost {
	std::vector<double> R;      /* kmax x kmax, column-major */
	std::vector<double> alpha;  /* accumulated coefficients of q_j */
};

template <typename T> static int ensure_inner(i3g_ctx *c, int m)
{
	if (c->in_m == m && c->in_prec == (int)sizeof(T)) return I3G_OK;
	for (void *q : c->iV) hipFree(q);
	for (void *q : c->iZ) hipFree(q);
	c->iV.clear(); c->iZ.clear();
	if (c->diV) hipFree(c->diV);
	if (c->diZ) hipFree(c->diZ);
	c->diV = c->diZ = nullptr;
	c->in_m = 0;
	long n = 4 * c->L[0].N;
	int err = 0;
	for (int q = 0; q <= m; q++) c->iV.push_back(dalloc<T>(n, &err));
	for (int q = 0; q < m; q++) c->iZ.push_back(dalloc<T>(n, &err));
	c->diV = dalloc<void *>(m + 1, &err);
	c->diZ = dalloc<void *>(m, &err);
	if (err) return err;
	if (upload(c->diV, c->iV.data(), m + 1) || upload(c->diZ, c->iZ.data(), m)) return I3G_EDEVICE;
	c->in_m = m;
	c->in_prec = (int)sizeof(T);
	return I3G_OK;
}

/* Inner FGMRES(m) in precision T on (W A) z = r with the V-cycle as right
 * preconditioner: at most m V-cycles, stopping early at a relative
 * reduction tol.  The operator product is evaluated in T (an FP32 inner
 * solve only has to reduce its residual by about tol, well above FP32
 * rounding).  out (T) receives z in the unscaled correction space.
 * Returns the V-cycles applied. */
template <typename T, typename R>
static int inner_fgmres(i3g_ctx *c, const R *r, T *out, int m, double tol, const i3g_stokes_params *p)
{
	Level &L = c->L[0];
	long N = L.N, n = 4 * N;
	dim3 lg = lin_grid(n);
	if (ensure_inner<T>(c, m)) return 0;
	T **V = (T **)c->iV.data(), **Z = (T **)c->iZ.data();
	T *const *dV = (T *const *)c->diV;
	double *dh = c->dh + 3 * c->kr_m + 4; /* 12 scratch doubles */
	std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
	hipLaunchKernelGGL((k_copy_scale<T, R>), lg, dim3(256), 0, c->s, V[0], r, 1.0, n);
	double beta = sqrt(dot(c, V[0], V[0], n));
	hipMemsetAsync(out, 0, n * sizeof(T), c->s);
	if (beta == 0) return 0;
	hipLaunchKernelGGL((k_copy_scale<T, T>), lg, dim3(256), 0, c->s, V[0], (const T *)V[0], 1.0 / beta, n);
	g[0] = beta;
	int j, used = 0;
	for (j = 0; j < m; j++) {
		T *x0 = lx<T>(L), *b0 = lb<T>(L);
		hipLaunchKernelGGL((k_wscale<T, T>), lg, dim3(256), 0, c->s, b0, (const T *)V[j], (const double *)L.wP64,
				   L.flags, N, 1, 1.0);
		hipMemsetAsync(x0, 0, n * sizeof(T), c->s);
		vcycle<T>(c, 0, p);
		used++;
		hipMemcpyAsync(Z[j], x0, n * sizeof(T), hipMemcpyDeviceToDevice, c->s);
		T *w = V[j + 1];
		{
			Timed tm(c, sizeof(T) == 8 ? "matvec_f64_L0" : "matvec_f32_L0", 0);
			hipLaunchKernelGGL((k_residual<T, T, 1>), plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s, view<T>(L),
					   (const T *)Z[j], (const T *)nullptr, w, L.bcid, bcview<T>(L), (double *)nullptr);
		}
		/* classical Gram-Schmidt; a second pass only when the first
		 * removed more than half of ||w||^2 (as in the outer GCR).  The
		 * first pass also returns ||w||^2 (w = V[j+1] is the last vector
		 * of the set). */
		double *h1 = c->dh, *h2 = c->dh + c->kr_m; /* kr_m >= m + 2 */
		multidot_dev<T, T>(c, (const T *const *)dV, j + 2, (const T *)w, n, h1);
		hipLaunchKernelGGL((k_multiaxpy<T, T>), lg, dim3(256), 0, c->s, w, (const T *const *)dV, (const double *)h1,
				   j + 1, n, -1.0);
		multidot_dev<T, T>(c, (const T *const *)(dV + j + 1), 1, (const T *)w, n, dh);
		hipMemcpyAsync(c->hscal + 16, h1, (j + 2) * sizeof(double), hipMemcpyDeviceToHost, c->s);
		hipMemcpyAsync(c->hscal + 8, dh, sizeof(double), hipMemcpyDeviceToHost, c->s);
		hipStreamSynchronize(c->s);
		for (int i = 0; i <= j; i++) H[(size_t)i * m + j] = c->hscal[16 + i];
		if (c->hscal[8] < 0.5 * c->hscal[16 + j + 1]) {
			multidot_dev<T, T>(c, (const T *const *)dV, j + 1, (const T *)w, n, h2);
			hipLaunchKernelGGL((k_multiaxpy<T, T>), lg, dim3(256), 0, c->s, w, (const T *const *)dV,
					   (const double *)h2, j + 1, n, -1.0);
			multidot_dev<T, T>(c, (const T *const *)(dV + j + 1), 1, (const T *)w, n, dh);
			hipMemcpyAsync(c->hscal + 16 + m + 2, h2, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, c->s);
			hipMemcpyAsync(c->hscal + 8, dh, sizeof(double), hipMemcpyDeviceToHost, c->s);
			hipStreamSynchronize(c->s);
			for (int i = 0; i <= j; i++) H[(size_t)i * m + j] += c->hscal[16 + m + 2 + i];
			c->reorth++;
		}
		hipLaunchKernelGGL(k_scale_rsqrt<T>, lg, dim3(256), 0, c->s, w, (const double *)dh, n);
		double hn = sqrt(c->hscal[8]);
		H[(size_t)(j + 1) * m + j] = hn;
		for (int i = 0; i < j; i++) {
			double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
			H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
			H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
		}
		double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
		double rr = sqrt(a * a + b * b);
		cs[j] = rr > 0 ? a / rr : 1.0;
		sn[j] = rr > 0 ? b / rr : 0.0;
		H[(size_t)j * m + j] = rr;
		H[(size_t)(j + 1) * m + j] = 0;
		g[j + 1] = -sn[j] * g[j];
		g[j] = cs[j] * g[j];
		if (fabs(g[j + 1]) <= tol * beta || hn == 0 || !std::isfinite(hn)) { j++; break; }
	}
	int k = j < m ? j : m;
	std::vector<double> y(k, 0.0);
	for (int i = k - 1; i >= 0; i--) {
		double s = g[i];
		for (int q = i + 1; q < k; q++) s -= H[(size_t)i * m + q] * y[q];
		y[i] = H[(size_t)i * m + i] != 0 ? s / H[(size_t)i * m + i] : 0;
	}
	hipMemcpyAsync(c->dh, y.data(), k * sizeof(double), hipMemcpyHostToDevice, c->s);
	hipLaunchKernelGGL((k_multiaxpy<T, T>), lg, dim3(256), 0, c->s, out, (const T *const *)c->diZ,
			   (const double *)c->dh, k, n, 1.0);
	hipStreamSynchronize(c->s);
	return used;
}

template <typename T, typename R> static void gcr_flush(i3g_ctx *c, GcrHost &G, int k, R *e)
{
	int m = c->kr_m;
	long n = 4 * c->L[0].N;
	if (k == 0) return;
	std::vector<double> y(k, 0);
	for (int i = k - 1; i >= 0; i--) {
		double s = G.alpha[i];
		for (int j = i + 1; j < k; j++) s -= G.R[(size_t)j * m + i] * y[j];
		y[i] = G.R[(size_t)i * m + i] != 0 ? s / G.R[(size_t)i * m + i] : 0;
	}
	hipMemcpyAsync(c->dh, y.data(), k * sizeof(double), hipMemcpyHostToDevice, c->s);
	hipLaunchKernelGGL((k_multiaxpy<R, T>), lin_grid(n), dim3(256), 0, c->s, e, (const T *const *)c->dP,
			   (const double *)c->dh, k, n, 1.0);
	hipStreamSynchronize(c->s);
	std::fill(G.alpha.begin(), G.alpha.end(), 0.0);
}

template <typename T, typename R>
static int gcr(i3g_ctx *c, GcrHost &G, R *r, R *e, double target, int max_new, const i3g_stokes_params *p,
	       double *red)
{
	Level &L = c->L[0];
	long N = L.N, n = 4 * N;
	int m = c->kr_m;
	T *const *dQ = (T *const *)c->dQ;
	double *h = c->dh + m, *h2 = c->dh + 2 * m, *sc = c->dh + 3 * m;
	dim3 lg = lin_grid(n);
	if (G.R.size() != (size_t)m * m) { G.R.assign((size_t)m * m, 0.0); G.alpha.assign(m, 0.0); }
	double r0 = sqrt(dot(c, r, r, n));
	if (red) *red = 1.0;
	if (r0 == 0) return 0;
	int k = c->kr_k;
	std::fill(G.alpha.begin(), G.alpha.end(), 0.0);
	if (k > 0) {
		/* recycle: remove the components of r in span(Q) */
		multidot_dev<T, R>(c, (const T *const *)dQ, k, (const R *)r, n, h);
		hipLaunchKernelGGL((k_multiaxpy<R, T>), lg, dim3(256), 0, c->s, r, (const T *const *)dQ, (const double *)h, k, n, -1.0);
		hipMemcpyAsync(c->hscal + 16, h, k * sizeof(double), hipMemcpyDeviceToHost, c->s);
		hipStreamSynchronize(c->s);
		for (int j = 0; j < k; j++) G.alpha[j] += c->hscal[16 + j];
	}
	int used = 0;
	double rn = sqrt(dot(c, r, r, n));
	/* the coupled smoother's guard: residual and V-cycles at the start of
	 * the current window */
	bool guard = c->L[0].scgs && !c->scgs_off && c->scgs_guard > 0;
	double rn_w = rn;
	int used_w = 0;
	while (rn > target * r0 && used < max_new) {
		if (k == m) {
			/* storage full: fold the directions into e and start afresh */
			gcr_flush<T, R>(c, G, k, e);
			k = 0;
		}
		T *pk = (T *)c->P[k], *qk = (T *)c->Q[k];
		if (p->inner_m > 0) {
			/* p_k = inner FGMRES(m) approximation of (W A)^-1 r */
			used += inner_fgmres<T, R>(c, r, pk, p->inner_m, p->inner_tol, p);
		} else {
			/* p_k = M W^-1 r */
			T *x0 = lx<T>(L), *b0 = lb<T>(L);
			hipLaunchKernelGGL((k_wscale<T, R>), lg, dim3(256), 0, c->s, b0, (const R *)r, (const double *)L.wP64, L.flags, N, 1, 1.0);
			hipMemsetAsync(x0, 0, n * sizeof(T), c->s);
			vcycle<T>(c, 0, p);
			used++;
			hipMemcpyAsync(pk, x0, n * sizeof(T), hipMemcpyDeviceToDevice, c->s);
		}
		/* q_k = W A p_k in FP64 */
		{
			Timed tm(c, "matvec_f64_L0", 0);
			hipLaunchKernelGGL((k_residual<double, T, 1>), plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s,
					   view<double>(L), (const T *)pk, (const double *)nullptr, L.r64, L.bcid,
					   bcview<double>(L), (double *)nullptr);
			hipLaunchKernelGGL((k_copy_scale<T, double>), lg, dim3(256), 0, c->s, qk, (const double *)L.r64, 1.0, n);
		}
		/* classical Gram-Schmidt against q_0..q_{k-1}; a second pass only
		 * when the first removed more than half of ||q||^2 */
		int again = 0;
		Timed tgs(c, "gcr_orthogonalise", 0);
		if (k > 0) {
			multidot_dev<T, T>(c, (const T *const *)(dQ + k), 1, (const T *)qk, n, sc + 3);
			multidot_dev<T, T>(c, (const T *const *)dQ, k, (const T *)qk, n, h);
			hipLaunchKernelGGL((k_multiaxpy<T, T>), lg, dim3(256), 0, c->s, qk, (const T *const *)dQ, (const double *)h, k, n, -1.0);
			hipMemcpyAsync(c->hscal + 16, h, k * sizeof(double), hipMemcpyDeviceToHost, c->s);
			hipMemcpyAsync(c->hscal + 8, sc + 3, sizeof(double), hipMemcpyDeviceToHost, c->s);
			hipStreamSynchronize(c->s);
			double s2 = 0;
			for (int j = 0; j < k; j++) s2 += c->hscal[16 + j] * c->hscal[16 + j];
			again = (c->hscal[8] - s2) < 0.5 * c->hscal[8];
			if (again) {
				multidot_dev<T, T>(c, (const T *const *)dQ, k, (const T *)qk, n, h2);
				hipLaunchKernelGGL((k_multiaxpy<T, T>), lg, dim3(256), 0, c->s, qk, (const T *const *)dQ, (const double *)h2, k, n, -1.0);
			}
		}
		multidot_dev<T, T>(c, (const T *const *)(dQ + k), 1, (const T *)qk, n, sc);
		hipLaunchKernelGGL(k_scale_rsqrt<T>, lg, dim3(256), 0, c->s, qk, (const double *)sc, n);
		multidot_dev<T, R>(c, (const T *const *)(dQ + k), 1, (const R *)r, n, sc + 1);
		hipLaunchKernelGGL((k_axpy_dev<R, T>), lg, dim3(256), 0, c->s, r, (const T *)qk, (const double *)(sc + 1), -1.0, n);
		{
			int nb = dot_blocks(n);
			hipLaunchKernelGGL((k_dot_part<R, R>), dim3(nb), dim3(256), 0, c->s, (const R *)r, (const R *)r, n, c->partials);
			hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->s, c->partials, (long)nb, 1, sc + 2);
		}
		/* one synchronisation per iteration: R column, alpha, ||r||^2 */
		hipMemcpyAsync(c->hscal + 16, h, 2 * m * sizeof(double) + 3 * sizeof(double), hipMemcpyDeviceToHost, c->s);
		hipStreamSynchronize(c->s);
		const double *hh = c->hscal + 16, *hh2 = hh + m, *scc = hh + 2 * m;
		for (int j = 0; j < k; j++) G.R[(size_t)k * m + j] = hh[j] + (again ? hh2[j] : 0.0);
		G.R[(size_t)k * m + k] = sqrt(scc[0]);
		G.alpha[k] = scc[1];
		rn = sqrt(scc[2]);
		k++;
		if (p->verbose > 1 || (p->verbose && used % 20 == 0))
			fprintf(stderr, "  gcr %d (v-cycles %d): rel %.3e\n", k, used, rn / r0);
		if (!std::isfinite(rn)) break;
		if (guard && used - used_w >= c->scgs_guard) {
			if (rn > c->scgs_guard_ratio * rn_w) {
				/* the coupled sweeps stall this operator: the colour
				 * phases from here on (GCR is flexible, the kept
				 * directions stay valid) */
				c->scgs_off = 1;
				guard = false;
				if (p->verbose)
					fprintf(stderr, "  gcr: coupled smoother off after %d V-cycles (rel %.3e over the last %d)\n",
						used, rn / rn_w, used - used_w);
			}
			rn_w = rn;
			used_w = used;
		}
	}
	gcr_flush<T, R>(c, G, k, e);
	c->kr_k = k;
	if (red) *red = rn / r0;
	return used;
}

template <typename T>
__global__ void k_add_scaled(double *__restrict__ x, const T *__restrict__ e, double a, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) x[t] += a * (double)e[t];
}

/* Pressure gauge: P += pinit - P(ref cell) (move3mg.c:2248-2269); the
 * shift is read first by a one-thread kernel, then applied. */
__global__ void k_pshift_get(const double *P, long ref, double pinit, double *shift)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) shift[0] = pinit - P[ref];
}
__global__ void k_pshift(double *P, long N, const double *shift, const uint8_t *flags)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < N && (flags[t] & FL_EX(0))) P[t] += shift[0];
}

extern "C" void i3g_stokes_default_params(i3g_stokes_params *p)
{
	memset(p, 0, sizeof(*p));
	p->rtol = 1e-8;
	p->precision = 32;
	p->max_outer = 60;
	p->krylov = 400;
	p->max_vcycles = 2000;
	/* Sweep counts: the reference's 8/16/16/32 per level and 64 on the
	 * coarsest (mode.t3c:3030-3031), with the finest and second level
	 * swapped to 16 and 8; the relaxation factors differ from mode.t3c
	 * (p0koef 0.3, p1koef 0.1, v0koef 0.7) because the V-cycle here
	 * preconditions a Krylov solver.  Measured on the 257^3 subduction case
	 * (10^6 contrast) with the V-cycle below: v0koef 0.75 / 0.85 / 0.95 /
	 * 1.0 / 1.05 took 597 / 447 / 406 / 406 / 419 V-cycles, p0koef
	 * 0.5 / 0.6 / 0.7 477 / 447 / 426, v1koef 0.6 / 0.7 / 0.85 487 / 447 /
	 * 448; v0koef 0.95, p0koef 0.7, v1koef 0.85 together 366 (its
	 * neighbours 356-387), and on 129^3 21 instead of 28 (32-blob cube) and
	 * 261 instead of 331 (subduction) against 0.85 / 0.6 / 0.7.
	 * Finest-level sweeps 8, 12, 16, 20, 24, 32 took 1538, 687, 477,
	 * 457, 447, 427 V-cycles, 16 being the fastest in time; 8 instead of 16
	 * sweeps on the second level cost 2% more V-cycles for 7% cheaper
	 * V-cycles.  The coarsest level (smoothed from LDS in one work-group)
	 * gets 128 sweeps: 457 instead of 477 V-cycles for 0.5 ms more per
	 * V-cycle; 256 gain nothing further.  8 instead of 16 sweeps on the
	 * third level kept the count (447) for 2.4% less time; 16 instead of
	 * 32 on the fourth and fifth cost 16% more V-cycles. */
	int pre[4] = {16, 8, 8, 32}, post[4] = {16, 8, 8, 32};
	for (int l = 0; l < I3G_MAXLEV; l++) {
		p->nu_pre[l] = l < 4 ? pre[l] : 32;
		p->nu_post[l] = l < 4 ? post[l] : 32;
	}
	p->nu_coarse = 128;
	p->inner_m = 10;
	p->inner_tol = 0.1;
	p->p0koef = 0.7; p->p1koef = 1.0; p->v0koef = 0.95; p->v1koef = 0.85;
	p->inner_rtol = 1e-6;
	p->warm_start = 0;
	p->pxinit = 2; p->pyinit = 1; p->pzinit = 2; p->pinit = 0;
}

static int solve_impl(i3g_ctx *c, const i3g_stokes_params *p, i3g_stats *st)
{
	Level &L = c->L[0];
	long N = L.N, n = 4 * N;
	i3g_stats loc;
	memset(&loc, 0, sizeof(loc));
	int m = p->krylov > 0 ? p->krylov : 1;
	int rc = p->precision == 64 ? ensure_krylov<double>(c, m) : ensure_krylov<float>(c, m);
	if (rc) return rc;
	m = c->kr_m;
	if (p->inner_m + 2 > m) return I3G_EINVAL; /* inner coefficients fit below the second-pass block */
	hipEvent_t t0, t1;
	hipEventCreate(&t0);
	hipEventCreate(&t1);
	hipEventRecord(t0, c->s);
	if (p->warm_start == 1) hipMemsetAsync(c->sol, 0, n * sizeof(double), c->s);
	bc_enforce<double>(c, L, c->sol, 0);
	double nrm[2];
	rc = resid64(c, c->sol, L.r64, nrm);
	if (rc) return rc;
	double bnorm = nrm[1] > 0 ? nrm[1] : 1.0;
	double rel = nrm[0] / bnorm;
	loc.res0 = rel;
	int outer = 0, cycles = 0;
	c->kr_k = 0; /* fresh direction set for every solve */
	GcrHost G;
	dim3 lg = lin_grid(n);
	while (rel > p->rtol && outer < p->max_outer) {
		double red = 1;
		double tgt = std::max(0.5 * p->rtol / rel, 1e-14);
		/* scaled, normalised FP64 residual; FP64 correction */
		double *r = (double *)c->kr, *e = (double *)c->ke;
		hipLaunchKernelGGL((k_wscale<double, double>), lg, dim3(256), 0, c->s, r, (const double *)L.r64,
				   (const double *)L.wP64, L.flags, N, 0, 1.0 / nrm[0]);
		hipMemsetAsync(e, 0, n * sizeof(double), c->s);
		if (p->precision == 64)
			cycles += gcr<double, double>(c, G, r, e, tgt, p->max_vcycles, p, &red);
		else {
			/* an FP32 pass reduces the scaled residual to about 1e-7;
			 * when that reaches the final target with margin, aim
			 * for it at once (a restart costs the kept directions'
			 * head start: 487 instead of 437 V-cycles from a start
			 * at 3.4e-2), else stop at inner_rtol and refine */
			double fin = 0.8 * p->rtol / rel;
			double itgt = fin >= 1e-7 ? fin : std::max(p->inner_rtol, tgt);
			cycles += gcr<float, double>(c, G, r, e, itgt, p->max_vcycles, p, &red);
		}
		hipLaunchKernelGGL(k_add_scaled<double>, lg, dim3(256), 0, c->s, c->sol, (const double *)e, nrm[0], n);
		bc_enforce<double>(c, L, c->sol, 0);
		rc = resid64(c, c->sol, L.r64, nrm);
		if (rc) return rc;
		rel = nrm[0] / bnorm;
		outer++;
		if (loc.nhist < 256) loc.hist[loc.nhist++] = rel;
		if (p->verbose)
			fprintf(stderr, "i3g outer %d: rel %.3e (inner est %.2e, cycles %d, second GS passes %ld)\n", outer,
				rel, red, cycles, c->reorth);
		if (!std::isfinite(rel)) break;
	}
	/* pressure gauge (FP64) */
	long ref = p->pzinit * L.nx * L.ny + p->pxinit * L.ny + p->pyinit;
	if (ref >= 0 && ref < N) {
		hipLaunchKernelGGL(k_pshift_get, dim3(1), dim3(64), 0, c->s, (const double *)c->sol, ref, p->pinit,
				   c->dscal + 8);
		hipLaunchKernelGGL(k_pshift, lin_grid(N), dim3(256), 0, c->s, c->sol, N, (const double *)(c->dscal + 8),
				   L.flags);
	}
	hipEventRecord(t1, c->s);
	HCHK(hipStreamSynchronize(c->s));
	float ms = 0;
	hipEventElapsedTime(&ms, t0, t1);
	hipEventDestroy(t0);
	hipEventDestroy(t1);
	flush_timing(c);
	loc.seconds = ms * 1e-3;
	loc.outer_iters = outer;
	loc.vcycles = cycles;
	loc.rel_res = rel;
	loc.converged = rel <= p->rtol;
	if (st) *st = loc;
	return loc.converged ? I3G_OK : I3G_ENOCONV;
}

/* unknowns that do not exist on the staggered grid are zero */
__global__ void k_zero_missing(double *x, long N, const uint8_t *flags)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= N) return;
	unsigned fl = flags[t];
	for (int q = 0; q < 4; q++)
		if (!(fl & FL_EX(q))) x[q * N + t] = 0.0;
}

extern "C" int i3g_stokes_set_solution(i3g_ctx *c, const double *P, const double *vx, const double *vy,
				       const double *vz)
{
	if (!c || !P || !vx || !vy || !vz) return I3G_EINVAL;
	Level &L = c->L[0];
	long N = L.N;
	const double *src[4] = {P, vx, vy, vz};
	for (int q = 0; q < 4; q++)
		HCHK(hipMemcpyAsync(c->sol + q * N, src[q], N * sizeof(double), hipMemcpyHostToDevice, c->s));
	/* the guess as the solver sees it: missing unknowns zero, boundary rows
	 * satisfied */
	hipLaunchKernelGGL(k_zero_missing, lin_grid(N), dim3(256), 0, c->s, c->sol, N, (const uint8_t *)L.flags);
	if (c->coef_ready) bc_enforce<double>(c, L, c->sol, 0);
	HCHK(hipGetLastError());
	HCHK(hipStreamSynchronize(c->s));
	return I3G_OK;
}

/* ---- level transfers on host planes (FP64) ----------------------------- */

static int planes_in(i3g_ctx *c, const double *const *pl, double *x, long N)
{
	for (int q = 0; q < 4; q++) {
		if (!pl[q]) return I3G_EINVAL;
		HCHK(hipMemcpyAsync(x + q * N, pl[q], N * sizeof(double), hipMemcpyHostToDevice, c->s));
	}
	return I3G_OK;
}

static int planes_out(i3g_ctx *c, double *const *pl, const double *x, long N)
{
	for (int q = 0; q < 4; q++) {
		if (!pl[q]) return I3G_EINVAL;
		HCHK(hipMemcpyAsync(pl[q], x + q * N, N * sizeof(double), hipMemcpyDeviceToHost, c->s));
	}
	HCHK(hipStreamSynchronize(c->s));
	return I3G_OK;
}

extern "C" int i3g_stokes_restrict(i3g_ctx *c, int level, const double *const *fine, double *const *coarse)
{
	if (!c || !c->coef_ready || !fine || !coarse || level < 1 || level >= c->nlev) return I3G_EINVAL;
	Level &F = c->L[level - 1], &Cl = c->L[level];
	int rc = planes_in(c, fine, F.r64, F.N);
	if (rc) return rc;
	hipLaunchKernelGGL(k_restrict<double>, comp_grid(Cl.nx, Cl.ny, Cl.nz), dim3(256), 0, c->s, view<double>(F),
			   view<double>(Cl), xview<double>(c->X[level - 1]), (const double *)F.r64, Cl.b64, Cl.x64);
	HCHK(hipGetLastError());
	return planes_out(c, coarse, Cl.b64, Cl.N);
}

extern "C" int i3g_stokes_prolong(i3g_ctx *c, int level, const double *const *coarse, double *const *fine,
				  double v1koef, double p1koef)
{
	if (!c || !c->coef_ready || !fine || !coarse || level < 1 || level >= c->nlev) return I3G_EINVAL;
	Level &F = c->L[level - 1], &Cl = c->L[level];
	int rc = planes_in(c, coarse, Cl.x64, Cl.N);
	if (!rc) rc = planes_in(c, (const double *const *)fine, F.x64, F.N);
	if (rc) return rc;
	hipLaunchKernelGGL(k_prolong<double>, plane_grid(F.nx, F.ny, F.nz), dim3(256), 0, c->s, view<double>(F),
			   view<double>(Cl), xview<double>(c->X[level - 1]), (const double *)Cl.x64, F.x64, v1koef,
			   p1koef);
	HCHK(hipGetLastError());
	return planes_out(c, fine, F.x64, F.N);
}

/* ---- the smoother on host planes (FP64 or rounded to FP32) -------------- */

/* x and b planes of level l into the level's work vectors of precision T */
template <typename T> static int smooth_in(i3g_ctx *c, Level &L, const double *const *x, const double *const *b)
{
	int rc = planes_in(c, (const double *const *)x, L.x64, L.N);
	if (!rc) rc = planes_in(c, b, L.b64, L.N);
	if (rc) return rc;
	if (sizeof(T) == 4) {
		long n = 4 * L.N;
		hipLaunchKernelGGL((k_copy_scale<float, double>), lin_grid(n), dim3(256), 0, c->s, L.x32,
				   (const double *)L.x64, 1.0, n);
		hipLaunchKernelGGL((k_copy_scale<float, double>), lin_grid(n), dim3(256), 0, c->s, L.b32,
				   (const double *)L.b64, 1.0, n);
	}
	return I3G_OK;
}

template <typename T> static int smooth_out(i3g_ctx *c, Level &L, double *const *x)
{
	if (sizeof(T) == 4) {
		long n = 4 * L.N;
		hipLaunchKernelGGL((k_copy_scale<double, float>), lin_grid(n), dim3(256), 0, c->s, L.x64,
				   (const float *)L.x32, 1.0, n);
	}
	HCHK(hipGetLastError());
	return planes_out(c, x, L.x64, L.N);
}

template <typename T>
static int smooth_phase_impl(i3g_ctx *c, int level, int phase, double *const *x, const double *const *b,
			     const i3g_stokes_params *p)
{
	Level &L = c->L[level];
	int rc = smooth_in<T>(c, L, x, b);
	if (rc) return rc;
	LevelView<T> v = view<T>(L);
	T *xx = lx<T>(L);
	const T *bb = lb<T>(L);
	if (phase == 6) {
		hipLaunchKernelGGL(k_smooth_p<T>, plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s, v, xx, bb,
				   (T)p->p0koef, 0);
	} else {
		dim3 hg = half_grid(L.nx, L.ny, L.nz);
		int chunks = (int)hg.x, col = phase & 1, nz = L.nz;
		T v0 = (T)p->v0koef;
		if (phase / 2 == 0)
			hipLaunchKernelGGL((k_smooth_vel_ilp<T, 1>), hg, dim3(256), 0, c->s, v, xx, bb, col, v0, 0, chunks, 0,
					   nz);
		else if (phase / 2 == 1)
			hipLaunchKernelGGL((k_smooth_vel_ilp<T, 2>), hg, dim3(256), 0, c->s, v, xx, bb, col, v0, 0, chunks, 0,
					   nz);
		else
			hipLaunchKernelGGL((k_smooth_vel_ilp<T, 3>), hg, dim3(256), 0, c->s, v, xx, bb, col, v0, 0, chunks, 0,
					   nz);
	}
	return smooth_out<T>(c, L, x);
}

extern "C" int i3g_stokes_smooth_phase(i3g_ctx *c, int level, int phase, int precision, double *const *x,
				       const double *const *b, double v0koef, double p0koef)
{
	if (!c || !c->coef_ready || !x || !b || level < 0 || level >= c->nlev || phase < 0 || phase > 6 ||
	    (precision != 32 && precision != 64))
		return I3G_EINVAL;
	i3g_stokes_params p;
	i3g_stokes_default_params(&p);
	p.v0koef = v0koef;
	p.p0koef = p0koef;
	return precision == 64 ? smooth_phase_impl<double>(c, level, phase, x, b, &p)
			       : smooth_phase_impl<float>(c, level, phase, x, b, &p);
}

template <typename T>
static int smooth_impl(i3g_ctx *c, int level, int nsweeps, int variant, double *const *x, const double *const *b,
		       const i3g_stokes_params *p)
{
	Level &L = c->L[level];
	int rc = smooth_in<T>(c, L, x, b);
	if (rc) return rc;
	int keep = c->fused, keepz = c->fused_minz, keeps = c->slab, keepa = c->slab_all;
	int keep7 = c->sweep7, keept = c->sweep7_min_tiles, keep3 = c->sweep3, keepg = c->sweep3_min_groups;
	if (variant >= 1 && variant <= 3) c->sweep7 = c->sweep3 = 0;
	if (variant == 4) c->sweep3 = 0;
	if (variant == 5) {
		c->sweep7 = 0;
		c->sweep3 = 1;
		c->sweep3_min_groups = 1;
		c->fused = 0;
		c->slab = 0;
	}
	if (variant == 1) { c->fused = 0; c->slab = 0; }
	if (variant == 2) { c->fused = 1; c->fused_minz = 0; }
	if (variant == 3) {
		c->fused = 0;
		c->slab = keeps > 0 ? keeps : 4;
		c->slab_all = 1;
	}
	if (variant == 4) {
		c->sweep7 = 1;
		c->sweep7_min_tiles = 1;
		c->fused = 0;
		c->slab = 0;
	}
	smooth_phases<T>(c, level, nsweeps, p);
	c->fused = keep;
	c->fused_minz = keepz;
	c->slab = keeps;
	c->slab_all = keepa;
	c->sweep7 = c->sweep7 ? keep7 : 0; /* a refused LDS opt-in stays off */
	c->sweep7_min_tiles = keept;
	c->sweep3 = c->sweep3 ? keep3 : 0;
	c->sweep3_min_groups = keepg;
	return smooth_out<T>(c, L, x);
}

extern "C" int i3g_stokes_smooth(i3g_ctx *c, int level, int precision, int nsweeps, int variant, double *const *x,
				 const double *const *b, double v0koef, double p0koef)
{
	if (!c || !c->coef_ready || !x || !b || level < 0 || level >= c->nlev || nsweeps < 0 || variant < 0 ||
	    variant > 5 || (precision != 32 && precision != 64))
		return I3G_EINVAL;
	i3g_stokes_params p;
	i3g_stokes_default_params(&p);
	p.v0koef = v0koef;
	p.p0koef = p0koef;
	return precision == 64 ? smooth_impl<double>(c, level, nsweeps, variant, x, b, &p)
			       : smooth_impl<float>(c, level, nsweeps, variant, x, b, &p);
}

/* host planes (NULL = zero) into a device 4N vector, missing unknowns
 * zeroed, boundary rows applied (homog: partners only) */
static int upload_planes(i3g_ctx *c, const double *const *pl, double *x, int homog,
			 hipMemcpyKind kind = hipMemcpyHostToDevice)
{
	Level &L = c->L[0];
	long N = L.N;
	for (int q = 0; q < 4; q++) {
		if (pl[q]) HCHK(hipMemcpyAsync(x + q * N, pl[q], N * sizeof(double), kind, c->s));
		else HCHK(hipMemsetAsync(x + q * N, 0, N * sizeof(double), c->s));
	}
	hipLaunchKernelGGL(k_zero_missing, lin_grid(N), dim3(256), 0, c->s, x, N, (const uint8_t *)L.flags);
	bc_enforce<double>(c, L, x, homog);
	HCHK(hipGetLastError());
	return I3G_OK;
}

static int combine_core(i3g_ctx *c, int k, const double *const *planes, hipMemcpyKind kind, double *a_out,
			double *rel_out)
{
	if (!c || !c->coef_ready || k < 0 || k > 4 || !planes) return I3G_EINVAL;
	Level &L = c->L[0];
	long N = L.N, n4 = 4 * N;
	int err = 0;
	if (c->cb_k < k) {
		/* grown once to the largest k used (the learned start uses 2) */
		if (c->cb_dirs) hipFree(c->cb_dirs);
		if (c->cb_ys) hipFree(c->cb_ys);
		c->cb_dirs = k ? dalloc<double>((size_t)k * n4, &err) : nullptr;
		c->cb_ys = dalloc<double>((size_t)(k + 1) * n4, &err);
		c->cb_k = err ? -1 : k;
		if (err) return err;
	}
	double *dirs = c->cb_dirs, *ys = c->cb_ys;
	int rc = upload_planes(c, planes, c->sol, 0, kind);
	double nrm[2] = {0, 1}, a[4] = {0, 0, 0, 0};
	if (!rc) rc = resid64(c, c->sol, L.r64, nrm);
	if (!rc) {
		dim3 lg = lin_grid(n4);
		/* y0 = W (b - A x0); y_i = W A d_i */
		hipLaunchKernelGGL((k_wscale<double, double>), lg, dim3(256), 0, c->s, ys, (const double *)L.r64,
				   (const double *)L.wP64, (const uint8_t *)L.flags, N, 0, 1.0);
		for (int i = 0; i < k && !rc; i++) {
			double *d = dirs + (long)i * n4;
			rc = upload_planes(c, planes + 4 * (i + 1), d, 1, kind);
			if (!rc)
				hipLaunchKernelGGL((k_residual<double, double, 1>), plane_grid(L.nx, L.ny, L.nz), dim3(256), 0,
						   c->s, view<double>(L), (const double *)d, (const double *)nullptr,
						   ys + (long)(i + 1) * n4, (const int *)L.bcid, bcview<double>(L),
						   (double *)nullptr);
		}
		if (!rc && k > 0) {
			/* normal equations G a = g of min || y0 - sum a_i y_i || */
			double G[4][5];
			for (int i = 0; i < k; i++) {
				for (int j = i; j < k; j++)
					G[i][j] = G[j][i] = dot(c, ys + (long)(i + 1) * n4, ys + (long)(j + 1) * n4, n4);
				G[i][k] = dot(c, ys, ys + (long)(i + 1) * n4, n4);
			}
			/* Gaussian elimination with partial pivoting */
			int ok = 1;
			for (int col = 0; col < k && ok; col++) {
				int piv = col;
				for (int r = col + 1; r < k; r++)
					if (fabs(G[r][col]) > fabs(G[piv][col])) piv = r;
				if (!(fabs(G[piv][col]) > 1e-300)) { ok = 0; break; }
				if (piv != col)
					for (int q = 0; q <= k; q++) std::swap(G[piv][q], G[col][q]);
				for (int r = col + 1; r < k; r++) {
					double f = G[r][col] / G[col][col];
					for (int q = col; q <= k; q++) G[r][q] -= f * G[col][q];
				}
			}
			if (ok)
				for (int i = k - 1; i >= 0; i--) {
					double v = G[i][k];
					for (int j = i + 1; j < k; j++) v -= G[i][j] * a[j];
					a[i] = v / G[i][i];
					if (!std::isfinite(a[i])) ok = 0;
				}
			if (!ok) a[0] = a[1] = a[2] = a[3] = 0;
			for (int i = 0; i < k; i++)
				if (a[i] != 0)
					hipLaunchKernelGGL((k_axpy<double, double>), lg, dim3(256), 0, c->s, c->sol,
							   (const double *)(dirs + (long)i * n4), a[i], n4);
			double nrm2[2];
			rc = resid64(c, c->sol, L.r64, nrm2);
			if (!rc && !(nrm2[0] <= nrm[0])) {
				/* rounding made it worse: back to the base */
				for (int i = 0; i < k; i++)
					if (a[i] != 0)
						hipLaunchKernelGGL((k_axpy<double, double>), lg, dim3(256), 0, c->s, c->sol,
								   (const double *)(dirs + (long)i * n4), -a[i], n4);
				a[0] = a[1] = a[2] = a[3] = 0;
			} else if (!rc) {
				nrm[0] = nrm2[0];
			}
		}
	}
	hipError_t e = hipStreamSynchronize(c->s);
	if (rc) return rc;
	if (e != hipSuccess) return I3G_EDEVICE;
	for (int i = 0; i < k; i++)
		if (a_out) a_out[i] = a[i];
	if (rel_out) *rel_out = nrm[0] / (nrm[1] > 0 ? nrm[1] : 1.0);
	return I3G_OK;
}

extern "C" int i3g_stokes_combine_start(i3g_ctx *c, int k, const double *const *planes, double *a_out,
					double *rel_out)
{
	return combine_core(c, k, planes, hipMemcpyHostToDevice, a_out, rel_out);
}

/* Learned start on the device.  Network features per node: (log10
 * max(nu, 1) - 21) / 2 and (rho - 3300) / 50. */
__global__ void k_ws_features(long N, const float *__restrict__ nu, const float *__restrict__ ro, float *f)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= N) return;
	double v = (double)nu[t];
	f[2 * t] = (float)((log10(v > 1.0 ? v : 1.0) - 21.0) / 2.0);
	f[2 * t + 1] = (float)(((double)ro[t] - 3300.0) / 50.0);
}

/* Lithostatic pressure, one thread per (m1, m3) column: the integral of
 * rho g_y down the column at the pressure points (cell centres, m1, m3 >= 1,
 * the cell's four corner densities in the x-z plane averaged; zero for
 * m2 < 2 and on the m1 = 0 and m3 = 0 columns). */
__global__ void k_ws_lithostatic(int nx, int ny, int nz, const float *__restrict__ ro, const double *gy, double g,
				 double *P)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nx * nz) return;
	int m3 = t / nx, m1 = t - m3 * nx;
	long sx = ny, sz = (long)nx * ny;
	long base = m3 * sz + m1 * sx;
	double acc = 0;
	P[base] = 0;
	if (ny > 1) P[base + 1] = 0;
	for (int m2 = 2; m2 < ny; m2++) {
		double rc = 0;
		if (m1 >= 1 && m3 >= 1) {
			long q = base + m2 - 1;
			rc = 0.25 * ((((double)ro[q] + (double)ro[q - sz]) + (double)ro[q - sx]) + (double)ro[q - sz - sx]);
		}
		double y1 = 0.5 * (gy[m2] + gy[m2 - 1]), y0 = m2 - 1 >= 1 ? 0.5 * (gy[m2 - 1] + gy[m2 - 2]) : 0.0;
		acc = acc + g * rc * (y1 - y0);
		P[base + m2] = acc;
	}
}

__global__ void k_ws_shift(long N, double *P, const double *ref)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < N) P[t] -= *ref;
}

__global__ void k_ws_copy1(const double *src, double *dst) { *dst = *src; }

/* scaled network outputs: plane q of out gets ch[q] * s[q] (channel
 * ch = 3 for P, 0..2 for vx..vz) */
__global__ void k_ws_outputs(long N, const float *__restrict__ y, double ps, double vs, double *dP, double *dv)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= N) return;
	dP[t] = (double)y[4 * t + 3] * ps;
	dv[t] = (double)y[4 * t] * vs;
	dv[N + t] = (double)y[4 * t + 1] * vs;
	dv[2 * N + t] = (double)y[4 * t + 2] * vs;
}

/* fallback 0: the zero start; 1: the resident iterate (the previous
 * step's solution in a time loop), kept in e64 while the others are tried */
extern "C" int i3g_stokes_learned_start_ex(i3g_ctx *c, i3g_net *net, double pscale, double vscale,
					   const long gauge[3], int fallback, double *info)
{
	if (!c || !c->coef_ready || !net || !gauge || !info || fallback < 0 || fallback > 1) return I3G_EINVAL;
	Level &L = c->L[0];
	long N = L.N;
	long g3 = gauge[2] * L.nx * L.ny + gauge[0] * L.ny + gauge[1];
	if (gauge[0] < 0 || gauge[0] >= L.nx || gauge[1] < 0 || gauge[1] >= L.ny || gauge[2] < 0 || gauge[2] >= L.nz)
		return I3G_EINVAL;
	int err = 0;
	if (!c->ws_feat) {
		c->ws_feat = dalloc<float>(2 * N, &err);
		c->ws_y = dalloc<float>(4 * N, &err);
		c->ws_lith = dalloc<double>(N + 1, &err);
		c->ws_dP = dalloc<double>(N, &err);
		c->ws_dv = dalloc<double>(3 * N, &err);
		if (err) return err;
	}
	float *feat = c->ws_feat, *y = c->ws_y;
	double *lith = c->ws_lith, *dP = c->ws_dP, *dv = c->ws_dv;
	int rc = I3G_OK;
	double relf = 1, rel_lith = 1, rel = 1, a[2] = {0, 0};
	int start = 0;
	const double *prev[4] = {c->e64, c->e64 + N, c->e64 + 2 * N, c->e64 + 3 * N};
	const double *zero[4] = {nullptr, nullptr, nullptr, nullptr};
	const double *const *fb = fallback ? prev : zero;
	if (fallback) {
		HCHK(hipMemcpyAsync(c->e64, c->sol, 4 * N * sizeof(double), hipMemcpyDeviceToDevice, c->s));
	}
	/* the fallback's residual */
	rc = combine_core(c, 0, fb, hipMemcpyDeviceToDevice, nullptr, &relf);
	if (!rc) {
		dim3 lg = lin_grid(N);
		hipLaunchKernelGGL(k_ws_features, lg, dim3(256), 0, c->s, N, (const float *)L.nu, (const float *)c->ro,
				   feat);
		hipLaunchKernelGGL(k_ws_lithostatic, dim3((unsigned)((L.nx * L.nz + 255) / 256)), dim3(256), 0, c->s,
				   L.nx, L.ny, L.nz, (const float *)c->ro, (const double *)L.g[1], c->phys[1], lith);
		hipLaunchKernelGGL(k_ws_copy1, dim3(1), dim3(1), 0, c->s, (const double *)(lith + g3), lith + N);
		hipLaunchKernelGGL(k_ws_shift, lg, dim3(256), 0, c->s, N, lith, (const double *)(lith + N));
		if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess) rc = I3G_EDEVICE;
	}
	if (!rc) rc = i3g_warmstart_infer_device(net, N, feat, y);
	if (!rc) {
		hipLaunchKernelGGL(k_ws_outputs, lin_grid(N), dim3(256), 0, c->s, N, (const float *)y, pscale, vscale, dP,
				   dv);
		const double *base[4] = {lith, nullptr, nullptr, nullptr};
		rc = combine_core(c, 0, base, hipMemcpyDeviceToDevice, nullptr, &rel_lith);
	}
	if (!rc) {
		const double *pl[12] = {lith, nullptr, nullptr, nullptr, dP, nullptr, nullptr, nullptr,
					nullptr, dv, dv + N, dv + 2 * N};
		rc = combine_core(c, 2, pl, hipMemcpyDeviceToDevice, a, &rel);
		start = 1;
		if (!rc && relf < rel) {
			rc = combine_core(c, 0, fb, hipMemcpyDeviceToDevice, nullptr, nullptr);
			start = 0;
		}
	}
	hipStreamSynchronize(c->s);
	if (rc) return rc;
	info[0] = a[0];
	info[1] = a[1];
	info[2] = rel;
	info[3] = rel_lith;
	info[4] = relf;
	info[5] = start;
	return I3G_OK;
}

extern "C" int i3g_stokes_learned_start(i3g_ctx *c, i3g_net *net, double pscale, double vscale,
					const long gauge[3], double *info)
{
	return i3g_stokes_learned_start_ex(c, net, pscale, vscale, gauge, 0, info);
}

/* velocity planes rounded to float and the FP64 pressure of the resident
 * solution into host or device arrays (any may be NULL), and the largest
 * |v| of each component over its existing unknowns (vmax may be NULL) */
__global__ void __launch_bounds__(256) k_fields(const double *__restrict__ sol, long N, const uint8_t *flags,
						float *vf, double *part)
{
	__shared__ double sh[3][4];
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double m[3] = {0, 0, 0};
	if (t < N) {
		unsigned fl = flags[t];
#pragma unroll
		for (int c = 0; c < 3; c++) {
			double v = sol[(c + 1) * N + t];
			vf[c * N + t] = (float)v;
			if (fl & FL_EX(c + 1)) m[c] = fabs((double)(float)v);
		}
	}
#pragma unroll
	for (int c = 0; c < 3; c++) {
		for (int o = 32; o > 0; o >>= 1) m[c] = fmax(m[c], __shfl_down(m[c], o, 64));
		if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = m[c];
	}
	__syncthreads();
	if (threadIdx.x < 3) {
		double a = 0;
		for (int w = 0; w < 4; w++) a = fmax(a, sh[threadIdx.x][w]);
		part[(long)blockIdx.x * 3 + threadIdx.x] = a;
	}
}

__global__ void k_fields_max(const double *part, long nb, double *out)
{
	if (threadIdx.x < 3) {
		double a = 0;
		for (long b = 0; b < nb; b++) a = fmax(a, part[b * 3 + threadIdx.x]);
		out[threadIdx.x] = a;
	}
}

extern "C" int i3g_stokes_fields(i3g_ctx *c, float *vx, float *vy, float *vz, double *pr, double *vmax)
{
	if (!c) return I3G_EINVAL;
	Level &L = c->L[0];
	long N = L.N;
	long nb = (N + 255) / 256;
	if (3 * nb > 2 * c->npart) {
		hipFree(c->partials);
		int err = 0;
		c->npart = (3 * nb + 1) / 2;
		c->partials = dalloc<double>(2 * c->npart, &err);
		if (err) return err;
	}
	hipLaunchKernelGGL(k_fields, dim3((unsigned)nb), dim3(256), 0, c->s, (const double *)c->sol, N,
			   (const uint8_t *)L.flags, L.r32, c->partials);
	hipLaunchKernelGGL(k_fields_max, dim3(1), dim3(64), 0, c->s, (const double *)c->partials, nb, c->dscal);
	HCHK(hipGetLastError());
	float *dst[3] = {vx, vy, vz};
	for (int q = 0; q < 3; q++)
		if (dst[q]) HCHK(hipMemcpyAsync(dst[q], L.r32 + q * N, N * sizeof(float), hipMemcpyDefault, c->s));
	if (pr) HCHK(hipMemcpyAsync(pr, c->sol, N * sizeof(double), hipMemcpyDefault, c->s));
	HCHK(hipMemcpyAsync(c->hscal, c->dscal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
	HCHK(hipStreamSynchronize(c->s));
	if (vmax)
		for (int q = 0; q < 3; q++) vmax[q] = c->hscal[q];
	return I3G_OK;
}

extern "C" int i3g_stokes_get_solution(i3g_ctx *c, double *P, double *vx, double *vy, double *vz)
{
	if (!c || !P || !vx || !vy || !vz) return I3G_EINVAL;
	long N = c->L[0].N;
	double *dst[4] = {P, vx, vy, vz};
	for (int q = 0; q < 4; q++)
		HCHK(hipMemcpyAsync(dst[q], c->sol + q * N, N * sizeof(double), hipMemcpyDeviceToHost, c->s));
	HCHK(hipStreamSynchronize(c->s));
	return I3G_OK;
}

extern "C" int i3g_stokes_solve_resident(i3g_ctx *c, const i3g_stokes_params *p, i3g_stats *st)
{
	if (!c || !p || !c->coef_ready) return I3G_EINVAL;
	if (p->precision != 32 && p->precision != 64) return I3G_EINVAL;
	for (int l = 0; l < c->nlev; l++)
		if (p->nu_pre[l] < 0 || p->nu_post[l] < 0) return I3G_EINVAL;
	return solve_impl(c, p, st);
}

extern "C" int i3g_stokes_solve(i3g_ctx *c, double *P, double *vx, double *vy, double *vz,
				const i3g_stokes_params *p, i3g_stats *st)
{
	if (!c || !p || !P || !vx || !vy || !vz) return I3G_EINVAL;
	int rc = i3g_stokes_set_solution(c, P, vx, vy, vz);
	if (rc) return rc;
	int rs = i3g_stokes_solve_resident(c, p, st);
	if (rs && rs != I3G_ENOCONV) return rs;
	rc = i3g_stokes_get_solution(c, P, vx, vy, vz);
	return rc ? rc : rs;
}

/* ------------------------------------------------------------------ */
/* Strain rates, stresses and the reference's native norms            */
/* ------------------------------------------------------------------ */

/* Value mode of sxxcalc..syzcalc (move3mg.c:3-710) over level 0, as the
 * loop at the end of vpiterate (:1691-1764): velocities rounded to float
 * (the reference's vx/vy/vz), differences in float, eps = diff * 1/h in
 * double, sigma = 2 eta eps, both stored as float.  st: 12 planes
 * exx eyy ezz exy exz eyz sxx syy szz sxy sxz syz; entries a component
 * does not have are left untouched. */
__global__ void k_strain_stress(LevelView<double> L, const double *__restrict__ sol, float *st)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / L.ny, m2 = t - m1 * L.ny;
	if (m1 >= L.nx) return;
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	long n = m3 * sz + m1 * sx + m2;
	const double *vx = sol + N, *vy = sol + 2 * N, *vz = sol + 3 * N;
#define F(a) ((float)(a))
	int nx = L.nx, ny = L.ny, nz = L.nz;
	if (m1 > 0 && m2 > 0 && m3 > 0) {
		long v0 = n - sz - sx - 1;
		double e;
		e = (double)(F(vx[v0 + sx]) - F(vx[v0])) * L.ihx[m1 - 1];
		st[n] = (float)e; st[6 * N + n] = (float)(2.0 * L.eC[n] * e);
		e = (double)(F(vy[v0 + 1]) - F(vy[v0])) * L.ihy[m2 - 1];
		st[N + n] = (float)e; st[7 * N + n] = (float)(2.0 * L.eC[n] * e);
		e = (double)(F(vz[v0 + sz]) - F(vz[v0])) * L.ihz[m3 - 1];
		st[2 * N + n] = (float)e; st[8 * N + n] = (float)(2.0 * L.eC[n] * e);
	}
	/* icx[m] = 1/(g[m+1]-g[m-1]) */
	if (m1 > 0 && m1 < nx - 1 && m2 > 0 && m2 < ny - 1 && m3 < nz - 1) {
		double e = (double)(F(vx[n]) - F(vx[n - 1])) * L.icy[m2] + (double)(F(vy[n]) - F(vy[n - sx])) * L.icx[m1];
		st[3 * N + n] = (float)e; st[9 * N + n] = (float)(2.0 * L.eXY[n] * e);
	}
	if (m1 > 0 && m1 < nx - 1 && m2 < ny - 1 && m3 > 0 && m3 < nz - 1) {
		double e = (double)(F(vx[n]) - F(vx[n - sz])) * L.icz[m3] + (double)(F(vz[n]) - F(vz[n - sx])) * L.icx[m1];
		st[4 * N + n] = (float)e; st[10 * N + n] = (float)(2.0 * L.eXZ[n] * e);
	}
	if (m1 < nx - 1 && m2 > 0 && m2 < ny - 1 && m3 > 0 && m3 < nz - 1) {
		double e = (double)(F(vy[n]) - F(vy[n - sz])) * L.icz[m3] + (double)(F(vz[n]) - F(vz[n - 1])) * L.icy[m2];
		st[5 * N + n] = (float)e; st[11 * N + n] = (float)(2.0 * L.eYZ[n] * e);
	}
#undef F
}

__device__ __forceinline__ double nueff48(const float *nu, long a0, long a1, long a2, long a3, long b0, long b1,
					  long b2, long b3, long c0, long c1, long c2, long c3, double nubeg,
					  double nuend)
{
	float s1 = nu[a0] + nu[a1] + nu[a2] + nu[a3];
	float s2 = nu[b0] + nu[b1] + nu[b2] + nu[b3];
	double e = ((double)s1 + 10.0 * (double)s2 + nu[c0] + nu[c1] + nu[c2] + nu[c3]) / 48.0;
	return e > nuend ? nuend : (e < nubeg ? nubeg : e);
}

/* Per-block sums of the norm loop of vpiterate (move3mg.c:2405-2560) on
 * level 0: momentum rows (x/y/zstokserr ynerr=1 from the stored float
 * stresses, double pressure, 48-weight viscosity for eps1[0]), continuity
 * rows (continerr ynerr=1) split by the P row's BC flag, velocity BC rows
 * (bonderr ynerr=1), and the float velocity ranges.  part[blk*16 + q]:
 * 0 sum ival^2, 1 sum eps^2, 2 sum eps^2 ival^2, 3 count, 4 cont^2,
 * 5 count, 6 contb^2, 7 count, 8 bond^2, 9 count, 10..15 min/max vx vy vz. */
#define NB_NORM 256
__global__ void __launch_bounds__(NB_NORM)
k_native_norms(LevelView<double> L, const double *__restrict__ sol, const float *__restrict__ st,
	       const float *__restrict__ nu, const float *__restrict__ ro, double g0, double g1, double g2,
	       double nubeg, double nuend, const int *__restrict__ bcid, BcRec<double> bc, double *part)
{
	double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
	double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / L.ny, m2 = t - m1 * L.ny;
	long N = L.N, sx = L.ny, sz = (long)L.nx * L.ny;
	int nx = L.nx, ny = L.ny, nz = L.nz;
	if (m1 < nx) {
		long n = m3 * sz + m1 * sx + m2;
		const double *P = sol;
		const float *sxx = st + 6 * N, *syy = st + 7 * N, *szz = st + 8 * N;
		const float *sxy = st + 9 * N, *sxz = st + 10 * N, *syz = st + 11 * N;
		unsigned fl = L.flags[n];
		if (m1 && m2 && m3) {
			long v0 = n - sz - sx - 1;
			const double *vx = sol + N, *vy = sol + 2 * N, *vz = sol + 3 * N;
			double xkf = L.ihx[m1 - 1], ykf = L.ihy[m2 - 1], zkf = L.ihz[m3 - 1];
			double lc = (double)((float)vx[v0 + sx] - (float)vx[v0]) * xkf +
				    (double)((float)vy[v0 + 1] - (float)vy[v0]) * ykf +
				    (double)((float)vz[v0 + sz] - (float)vz[v0]) * zkf;
			double skf = (1.0 / xkf + 1.0 / ykf + 1.0 / zkf) / 3.0;
			lc = (lc - 0) * skf;
			if (!(fl & FL_BC(0))) { acc[4] += lc * lc; acc[5] += 1; }
			else { acc[6] += lc * lc; acc[7] += 1; }
		}
		for (int c = 1; c <= 3; c++) {
			if (!(fl & FL_EX(c))) continue;
			double vv = (double)(float)sol[(long)c * N + n];
			mn[c - 1] = fmin(mn[c - 1], vv);
			mx[c - 1] = fmax(mx[c - 1], vv);
			if (fl & FL_BC(c)) {
				int r = bcid[(long)c * N + n];
				double lb = sol[(long)c * N + n] - bc.c[r];
				if (bc.partner[r] >= 0) lb -= bc.k[r] * sol[bc.partner[r]];
				acc[8] += lb * lb; acc[9] += 1;
				continue;
			}
			long v[8] = {n, n + 1, n + sx, n + sx + 1, n + sz, n + sz + 1, n + sz + sx, n + sz + sx + 1};
			double xkf, ykf, zkf, right, left, ne;
			if (c == 1) {
				xkf = 2.0 * L.icx[m1]; ykf = L.ihy[m2]; zkf = L.ihz[m3];
				right = -g0 * (double)(ro[v[0]] + ro[v[1]] + ro[v[4]] + ro[v[5]]) / 4.0;
				ne = nueff48(nu, v[0] - sx, v[1] - sx, v[4] - sx, v[5] - sx, v[0], v[1], v[4], v[5], v[2],
					     v[3], v[6], v[7], nubeg, nuend);
				left = (double)(sxx[v[7]] - sxx[v[5]]) * xkf;
				left += (double)(sxy[v[1]] - sxy[v[0]]) * ykf;
				left += (double)(sxz[v[4]] - sxz[v[0]]) * zkf;
				left -= (P[v[7]] - P[v[5]]) * xkf;
			} else if (c == 2) {
				xkf = L.ihx[m1]; ykf = 2.0 * L.icy[m2]; zkf = L.ihz[m3];
				right = -g1 * (double)(ro[v[0]] + ro[v[2]] + ro[v[4]] + ro[v[6]]) / 4.0;
				ne = nueff48(nu, v[0] - 1, v[2] - 1, v[4] - 1, v[6] - 1, v[0], v[2], v[4], v[6], v[1], v[3],
					     v[5], v[7], nubeg, nuend);
				left = (double)(syy[v[7]] - syy[v[6]]) * ykf;
				left += (double)(sxy[v[2]] - sxy[v[0]]) * xkf;
				left += (double)(syz[v[4]] - syz[v[0]]) * zkf;
				left -= (P[v[7]] - P[v[6]]) * ykf;
			} else {
				xkf = L.ihx[m1]; ykf = L.ihy[m2]; zkf = 2.0 * L.icz[m3];
				right = -g2 * (double)(ro[v[0]] + ro[v[1]] + ro[v[2]] + ro[v[3]]) / 4.0;
				ne = nueff48(nu, v[0] - sz, v[1] - sz, v[2] - sz, v[3] - sz, v[0], v[1], v[2], v[3], v[4],
					     v[5], v[6], v[7], nubeg, nuend);
				left = (double)(szz[v[7]] - szz[v[3]]) * zkf;
				left += (double)(sxz[v[2]] - sxz[v[0]]) * xkf;
				left += (double)(syz[v[1]] - syz[v[0]]) * ykf;
				left -= (P[v[7]] - P[v[3]]) * zkf;
			}
			double skf1 = 1.0 / ne * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
			left = left - right;
			double e = left * skf1;
			acc[0] += left * left; acc[1] += e * e; acc[2] += e * e * left * left; acc[3] += 1;
		}
	}
	__shared__ double sh[NB_NORM / 64][16];
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	for (int q = 0; q < 10; q++) {
		double v = acc[q];
		for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
		if (lane == 0) sh[w][q] = v;
	}
	for (int q = 0; q < 3; q++) {
		double a = mn[q], b = mx[q];
		for (int o = 32; o > 0; o >>= 1) {
			a = fmin(a, __shfl_down(a, o, 64));
			b = fmax(b, __shfl_down(b, o, 64));
		}
		if (lane == 0) { sh[w][10 + 2 * q] = a; sh[w][11 + 2 * q] = b; }
	}
	__syncthreads();
	if (threadIdx.x < 16) {
		int q = threadIdx.x;
		double v = sh[0][q];
		for (int k = 1; k < NB_NORM / 64; k++) {
			double u = sh[k][q];
			if (q < 10) v += u;
			else if ((q - 10) % 2 == 0) v = fmin(v, u);
			else v = fmax(v, u);
		}
		long blk = (long)blockIdx.y * gridDim.x + blockIdx.x;
		part[blk * 16 + q] = v;
	}
}

__global__ void k_native_final(long nblk, const double *part, double *out)
{
	__shared__ double sh[256][16];
	double v[16];
	for (int q = 0; q < 16; q++) v[q] = q < 10 ? 0.0 : ((q - 10) % 2 == 0 ? 1e300 : -1e300);
	for (long b = threadIdx.x; b < nblk; b += 256)
		for (int q = 0; q < 16; q++) {
			double u = part[b * 16 + q];
			if (q < 10) v[q] += u;
			else if ((q - 10) % 2 == 0) v[q] = fmin(v[q], u);
			else v[q] = fmax(v[q], u);
		}
	for (int q = 0; q < 16; q++) sh[threadIdx.x][q] = v[q];
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o)
			for (int q = 0; q < 16; q++) {
				double a = sh[threadIdx.x][q], b = sh[threadIdx.x + o][q];
				sh[threadIdx.x][q] = q < 10 ? a + b : ((q - 10) % 2 == 0 ? fmin(a, b) : fmax(a, b));
			}
		__syncthreads();
	}
	if (threadIdx.x < 16) out[threadIdx.x] = sh[0][threadIdx.x];
}

static int ensure_st(i3g_ctx *c)
{
	if (c->st12) return I3G_OK;
	int err = 0;
	c->st12 = dalloc<float>(12 * c->L[0].N, &err);
	if (err) return err;
	HCHK(hipMemsetAsync(c->st12, 0, 12 * c->L[0].N * sizeof(float), c->s));
	return I3G_OK;
}

extern "C" int i3g_stokes_strain_stress(i3g_ctx *c, float *exx, float *eyy, float *ezz, float *exy, float *exz,
					float *eyz, float *sxx, float *syy, float *szz, float *sxy, float *sxz,
					float *syz)
{
	if (!c || !c->coef_ready) return I3G_EINVAL;
	int rc = ensure_st(c);
	if (rc) return rc;
	Level &L = c->L[0];
	hipLaunchKernelGGL(k_strain_stress, plane_grid(L.nx, L.ny, L.nz), dim3(256), 0, c->s, view<double>(L),
			   (const double *)c->sol, c->st12);
	HCHK(hipGetLastError());
	float *dst[12] = {exx, eyy, ezz, exy, exz, eyz, sxx, syy, szz, sxy, sxz, syz};
	for (int q = 0; q < 12; q++)
		if (dst[q])
			HCHK(hipMemcpyAsync(dst[q], c->st12 + (long)q * L.N, L.N * sizeof(float), hipMemcpyDeviceToHost,
					    c->s));
	HCHK(hipStreamSynchronize(c->s));
	return I3G_OK;
}

extern "C" int i3g_stokes_native_norms(i3g_ctx *c, double divvmax, double *out)
{
	if (!c || !out || !c->coef_ready) return I3G_EINVAL;
	int rc = ensure_st(c);
	if (rc) return rc;
	Level &L = c->L[0];
	dim3 pg = plane_grid(L.nx, L.ny, L.nz);
	long nblk = (long)pg.x * pg.y;
	hipLaunchKernelGGL(k_strain_stress, pg, dim3(256), 0, c->s, view<double>(L), (const double *)c->sol, c->st12);
	double *part = nullptr;
	int err = 0;
	part = dalloc<double>(16 * nblk + 16, &err);
	if (err) return err;
	hipLaunchKernelGGL(k_native_norms, pg, dim3(NB_NORM), 0, c->s, view<double>(L), (const double *)c->sol,
			   (const float *)c->st12, (const float *)L.nu, (const float *)c->ro, c->phys[0], c->phys[1],
			   c->phys[2], c->phys[3], c->phys[4], (const int *)L.bcid, bcview<double>(L), part);
	hipLaunchKernelGGL(k_native_final, dim3(1), dim3(256), 0, c->s, nblk, (const double *)part, part + 16 * nblk);
	double h[16];
	hipError_t e1 = hipGetLastError();
	hipError_t e2 = hipMemcpyAsync(h, part + 16 * nblk, 16 * sizeof(double), hipMemcpyDeviceToHost, c->s);
	hipError_t e3 = hipStreamSynchronize(c->s);
	hipFree(part);
	if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) return I3G_EDEVICE;
	/* maxvxyz (move3mg.c:1766-1770) */
	double dvx = h[11] - h[10], dvy = h[13] - h[12], dvz = h[15] - h[14];
	double mv = sqrt((dvx * dvx + dvy * dvy + dvz * dvz) / 3.0);
	if (mv <= 0) mv = 1e-8;
	if (divvmax > 0 && mv > divvmax) mv = divvmax;
	out[0] = sqrt(h[0] / h[3]);
	out[1] = sqrt(h[1] / h[3]) / mv;
	out[2] = sqrt(h[2] / h[3]) / mv;
	out[3] = h[5] > 0 ? sqrt(h[4] / h[5]) / mv : 0;
	out[4] = h[7] > 0 ? sqrt(h[6] / h[7]) / mv : 0;
	out[5] = h[9] > 0 ? sqrt(h[8] / h[9]) : 0;
	out[6] = mv;
	return I3G_OK;
}

extern "C" int i3g_set_timing(i3g_ctx *c, int on)
{
	if (!c) return I3G_EINVAL;
	c->timing = on != 0;
	c->stats.clear();
	return I3G_OK;
}

extern "C" int i3g_kernel_stats(i3g_ctx *c, char *names, int name_len, double *avg_ms, long *calls,
				double *bytes, int cap)
{
	if (!c) return I3G_EINVAL;
	flush_timing(c);
	int q = 0;
	for (auto &kv : c->stats) {
		if (q >= cap) break;
		if (names) {
			strncpy(names + (size_t)q * name_len, kv.first.c_str(), name_len - 1);
			names[(size_t)q * name_len + name_len - 1] = 0;
		}
		if (avg_ms) avg_ms[q] = kv.second.calls ? kv.second.ms / kv.second.calls : 0;
		if (calls) calls[q] = kv.second.calls;
		if (bytes) bytes[q] = kv.second.calls ? kv.second.bytes / kv.second.calls : 0;
		q++;
	}
	return q;
}
