/*
 * i3g markers on MI355X: cell search, deterministic cell sort, marker ->
 * node transfer (gather form of the reference's ronurecalc scatter) and RK4
 * advection.  C ABI in include/i3g.h.
 */
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/i3g.h"

/* The reference evaluates every product and sum with its own rounding (no
 * fused multiply-add on its x86-64 build); keep that for bit parity. */
#pragma clang fp contract(off)

#include "grid_interp.h"
#include "rheology.h"

#define MCHK(x)                                                                \
	do {                                                                   \
		hipError_t e_ = (x);                                           \
		if (e_ != hipSuccess) {                                        \
			fprintf(stderr, "i3g markers: HIP error %s at %s:%d\n", \
				hipGetErrorString(e_), __FILE__, __LINE__);    \
			return I3G_EDEVICE;                                    \
		}                                                              \
	} while (0)

#define MAXROCK 64

struct i3g_markers {
	hipStream_t s = nullptr;
	int nx = 0, ny = 0, nz = 0;
	long N = 0;              /* nodes */
	long cap = 0, n = 0;     /* marker capacity / count */
	double *gx = nullptr, *gy = nullptr, *gz = nullptr;
	/* SoA marker state (current order) and a second buffer for permuting */
	float *x = nullptr, *y = nullptr, *z = nullptr, *tk = nullptr, *eta = nullptr;
	float *e = nullptr;      /* accumulated strain (marke) */
	float *rr = nullptr, *rn = nullptr; /* grain roughness (markrr), dislocation share (markrn) */
	signed char *type = nullptr;
	int *id = nullptr;
	float *x2 = nullptr, *y2 = nullptr, *z2 = nullptr, *tk2 = nullptr, *eta2 = nullptr, *e2 = nullptr;
	float *rr2 = nullptr, *rn2 = nullptr;
	signed char *type2 = nullptr;
	int *id2 = nullptr;
	int *key = nullptr, *c1 = nullptr, *c2 = nullptr, *c3 = nullptr;
	int *perm = nullptr;     /* sorted position -> previous position */
	/* cell bins over nodes: start[N+1], count/cursor[N] */
	int *cnt = nullptr, *start = nullptr, *cur = nullptr;
	int *bsum = nullptr;     /* scan block sums */
	int sorted = 0;
	/* per-marker records for the node transfer (sorted order) */
	struct MkRec *prec = nullptr;
	/* node-field staging */
	double *pr = nullptr;
	float *fld = nullptr;    /* 11 x N float in/out fields */
	float *vel = nullptr;    /* 3 x N velocities */
	float *st12 = nullptr;   /* 12 x N strain rates and stresses (epscalc1) */
	int has_st = 0;          /* st12 holds the planes of i3g_markers_set_stress */
	float *md = nullptr;     /* marker density by id (fluid velocity) */
	long md_cap = 0;
	/* water content (markw) and extracted melt (markex) by id, of the melt
	 * types' properties; mw_n ids, none until i3g_markers_set_melt */
	float *mw = nullptr;
	long mw_n = 0, mw_cap = 0;
	/* node averages: 1 cell-segmented (k_cell_transfer / k_node_sums), 0 the
	 * reference's serial marker order (k_marker_gather, k_tk_gather) */
	int order = 1;
	/* 1: the cell-segmented transfer stages its records in prec
	 * (k_marker_props + k_cell_partials) instead of computing them in the
	 * reduction (I3G_MK_STAGED=1 at create); the same sums to the bit */
	int staged = 0;
	double *part = nullptr; /* per-cell corner partials, PF_N x 8 x cells */
	/* movemark's depth rules (inactive by default) */
	double zdeep = 1e30, vdeep = 1e30, tdeep = 0;
	i3g_rock *rocks = nullptr;
	int *bad = nullptr;
	/* temperature boundary rows of level 0 (tkrecalc): per node partner
	 * (-2 none, -1 constant only), constant and coefficient */
	int *tbc_p = nullptr;
	float *tbc_c = nullptr, *tbc_k = nullptr, *tk1 = nullptr;
	int has_tbc = 0;
	/* marker temperature update: node fields tk, tk2, kt, ro, cp (5 x N),
	 * per-marker diffusion change and flag, averaged nodal change */
	float *tnod = nullptr;
	double *sn = nullptr;
	struct TkRec *trec = nullptr; /* per-marker records for the nodal average (sorted order) */
	uint8_t *tin = nullptr;
	/* device time of the last operation's kernels (host copies excluded) */
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	float last_ms = 0;
};

static void mark_begin(i3g_markers *m) { hipEventRecord(m->ev0, m->s); }
static void mark_end(i3g_markers *m) { hipEventRecord(m->ev1, m->s); }
/* after a synchronisation of the stream */
static void mark_read(i3g_markers *m)
{
	float ms = 0;
	if (hipEventElapsedTime(&ms, m->ev0, m->ev1) == hipSuccess) m->last_ms = ms;
}

/* ------------------------------------------------------------------ */
/* Cell search                                                        */
/* ------------------------------------------------------------------ */

/* The grid lines of the three axes staged in LDS (dynamic shared memory of
 * nx+ny+nz doubles): every marker's binary searches and interpolation
 * weights read them many times at lane-divergent addresses.  Called by all
 * threads of the group before any early exit. */
__device__ __forceinline__ void stage_lines(double *sg, const double *gx, const double *gy, const double *gz, int nx,
					    int ny, int nz)
{
	for (int i = threadIdx.x; i < nx + ny + nz; i += blockDim.x)
		sg[i] = i < nx ? gx[i] : (i < nx + ny ? gy[i - nx] : gz[i - nx - ny]);
	__syncthreads();
}

__global__ void k_cell_index(long n, const float *__restrict__ x, const float *__restrict__ y,
			     const float *__restrict__ z, const double *gx, const double *gy, const double *gz,
			     int nx, int ny, int nz, int *key, int *c1, int *c2, int *c3, int *cnt)
{
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	double a = x[t], b = y[t], c = z[t];
	if (a < 0) a = 0;
	if (a > xs) a = xs;
	if (b < 0) b = 0;
	if (b > ys) b = ys;
	if (c < 0) c = 0;
	if (c > zs) c = zs;
	int m1 = serch(gx, nx, a), m2 = serch(gy, ny, b), m3 = serch(gz, nz, c);
	int k = m3 * nx * ny + m1 * ny + m2;
	key[t] = k;
	if (c1) { c1[t] = m1; c2[t] = m2; c3[t] = m3; }
	if (cnt) atomicAdd(&cnt[k], 1);
}

/* ------------------------------------------------------------------ */
/* Counting sort by cell, then by id within a cell                    */
/* ------------------------------------------------------------------ */

#define SCAN_B 1024

/* exclusive scan of cnt[0..N) into start[0..N], three passes */
__global__ void k_scan_local(const int *cnt, int *start, long N, int *bsum)
{
	__shared__ int s[SCAN_B];
	long base = (long)blockIdx.x * SCAN_B;
	int t = threadIdx.x;
	int v = base + t < N ? cnt[base + t] : 0;
	s[t] = v;
	__syncthreads();
	for (int off = 1; off < SCAN_B; off <<= 1) {
		int a = t >= off ? s[t - off] : 0;
		__syncthreads();
		s[t] += a;
		__syncthreads();
	}
	if (base + t < N) start[base + t] = s[t] - v;
	if (t == SCAN_B - 1) bsum[blockIdx.x] = s[t];
}

__global__ void k_scan_bsum(int *bsum, long nb)
{
	/* one block, sequential over chunks of 1024 */
	__shared__ int s[SCAN_B];
	__shared__ int carry;
	if (threadIdx.x == 0) carry = 0;
	__syncthreads();
	for (long base = 0; base < nb; base += SCAN_B) {
		int t = threadIdx.x;
		int v = base + t < nb ? bsum[base + t] : 0;
		s[t] = v;
		__syncthreads();
		for (int off = 1; off < SCAN_B; off <<= 1) {
			int a = t >= off ? s[t - off] : 0;
			__syncthreads();
			s[t] += a;
			__syncthreads();
		}
		if (base + t < nb) bsum[base + t] = s[t] - v + carry;
		__syncthreads();
		if (t == SCAN_B - 1) carry += s[t];
		__syncthreads();
	}
}

__global__ void k_scan_add(int *start, long N, const int *bsum, long total_slot)
{
	long i = (long)blockIdx.x * SCAN_B + threadIdx.x;
	if (i < N) start[i] += bsum[blockIdx.x];
	(void)total_slot;
}

/* A marker's whole state in the sort's staging buffer: 48 bytes, three
 * 16-byte words (position and temperature; viscosity, strain, roughness,
 * dislocation share; type, id, previous position). */
struct SortRec {
	float4 a, b;
	int4 c;
};

struct MkArrays {
	float *x, *y, *z, *tk, *eta, *e, *rr, *rn;
	signed char *type;
	int *id;
};

/* Each marker's record into its cell's range (in arrival order).  Lanes of
 * a wavefront with the same cell in a row (markers that are already sorted
 * arrive that way) take their slots with one atomic per run. */
__global__ void __launch_bounds__(256) k_fill(long n, const int *__restrict__ key, const int *__restrict__ start,
					      int *cur, MkArrays src, SortRec *__restrict__ rec)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	const int lane = threadIdx.x & 63;
	const bool act = t < n;
	int k = act ? key[t] : -1;
	int prev = __shfl_up(k, 1);
	bool head = act && (lane == 0 || prev != k);
	unsigned long long heads = __ballot(head), acts = __ballot(act);
	unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
	int rs = 63 - __clzll(heads & upto); /* first lane of this lane's run */
	unsigned long long after = heads & ~upto;
	int re = after ? __ffsll(after) - 1 : 64 - __clzll(acts); /* one past its last lane */
	int base = 0;
	if (head) base = atomicAdd(&cur[k], re - lane);
	base = __shfl(base, rs);
	if (!act) return;
	int pos = start[k] + base + (lane - rs);
	SortRec r;
	r.a = make_float4(src.x[t], src.y[t], src.z[t], src.tk[t]);
	r.b = make_float4(src.eta[t], src.e[t], src.rr[t], src.rn[t]);
	r.c = make_int4((int)src.type[t], src.id[t], (int)t, 0);
	rec[pos] = r;
}

/* Within each cell, order the markers by id and write them out: one
 * wavefront per cell sorts the cell's (id, slot) pairs with a bitonic
 * network (in registers across the lanes for up to 64 markers, in LDS for
 * up to CS_MAX), then writes every marker's arrays, its previous position
 * and its cell indices in the sorted order.  A cell of more than CS_MAX
 * markers is sorted in chunks of CS_MAX in LDS and the sorted chunks are
 * merged pairwise in `scratch` (merge path: each lane merges one 64th of
 * the output after a binary search on its diagonal), so its cost grows as
 * e log e, not e^2.  Ids are unique, so the order is fully determined (the
 * same as any stable sort by id). */
#define CS_MAX 1024
__device__ __forceinline__ void put_sorted(const SortRec &r, long o, int c, int nx, int ny, MkArrays d, int *perm,
					   int *key, int *c1, int *c2, int *c3)
{
	d.x[o] = r.a.x; d.y[o] = r.a.y; d.z[o] = r.a.z; d.tk[o] = r.a.w;
	d.eta[o] = r.b.x; d.e[o] = r.b.y; d.rr[o] = r.b.z; d.rn[o] = r.b.w;
	d.type[o] = (signed char)r.c.x;
	d.id[o] = r.c.y;
	perm[o] = r.c.z;
	int pl = nx * ny, m3 = c / pl, rem = c - m3 * pl, m1 = rem / ny;
	key[o] = c;
	c1[o] = m1;
	c2[o] = rem - m1 * ny;
	c3[o] = m3;
}

/* Bitonic sort of the e <= CS_MAX records rec[0..e) by id into K (ids) and
 * V (indices into rec) in one wavefront's LDS. */
__device__ __forceinline__ void lds_bitonic(const SortRec *__restrict__ rec, int e, int *K, int *V, int lane)
{
	int P = 128;
	while (P < e) P <<= 1;
	for (int i = lane; i < P; i += 64) {
		K[i] = i < e ? rec[i].c.y : INT_MAX;
		V[i] = i;
	}
	/* one wavefront owns K, V: its LDS operations complete in program
	 * order; the empty asm keeps the compiler from reordering them */
	asm volatile("" ::: "memory");
	for (int kk = 2; kk <= P; kk <<= 1)
		for (int j = kk >> 1; j > 0; j >>= 1) {
			for (int i = lane; i < P; i += 64) {
				int ixj = i ^ j;
				if (ixj > i) {
					int a = K[i], b = K[ixj];
					if ((a > b) == ((i & kk) == 0)) {
						int va = V[i];
						K[i] = b;
						K[ixj] = a;
						V[i] = V[ixj];
						V[ixj] = va;
					}
				}
			}
			asm volatile("" ::: "memory");
		}
}

/* Merge the sorted runs (ka, va)[0..na) and (kb, vb)[0..nb) (distinct
 * keys) into (ko, vo)[0..na+nb): lane l writes outputs [l*per, (l+1)*per),
 * starting where the merge path crosses that diagonal. */
__device__ __forceinline__ void wave_merge(const int *ka, const int *va, int na, const int *kb, const int *vb, int nb,
					   int *ko, int *vo, int lane)
{
	const int tot = na + nb, per = (tot + 63) / 64;
	const int d0 = min(lane * per, tot), d1 = min(d0 + per, tot);
	if (d0 >= d1) return;
	int lo = max(0, d0 - nb), hi = min(d0, na);
	while (lo < hi) {
		int mid = (lo + hi) >> 1;
		if (ka[mid] < kb[d0 - 1 - mid]) lo = mid + 1;
		else hi = mid;
	}
	int i = lo, j = d0 - lo;
	int a = i < na ? ka[i] : INT_MAX, b = j < nb ? kb[j] : INT_MAX;
	for (int o = d0; o < d1; o++) {
		if (i < na && (j >= nb || a < b)) {
			ko[o] = a;
			vo[o] = va[i];
			i++;
			a = i < na ? ka[i] : INT_MAX;
		} else {
			ko[o] = b;
			vo[o] = vb[j];
			j++;
			b = j < nb ? kb[j] : INT_MAX;
		}
	}
}

__global__ void __launch_bounds__(256) k_cellsort(long N, long cap, int nx, int ny, const int *__restrict__ start,
						  const int *__restrict__ cnt, const SortRec *__restrict__ rec, MkArrays d,
						  int *perm, int *key, int *c1, int *c2, int *c3, int *scratch)
{
	__shared__ int sk[4][CS_MAX], sv[4][CS_MAX];
	const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
	long cl = (long)blockIdx.x * 4 + w;
	if (cl >= N) return;
	const int c = (int)cl, s = start[c], e = cnt[c];
	if (e == 0) return;
	if (e <= 64) {
		int P = 1;
		while (P < e) P <<= 1;
		int k = lane < e ? rec[s + lane].c.y : INT_MAX, v = lane;
		for (int kk = 2; kk <= P; kk <<= 1)
			for (int j = kk >> 1; j > 0; j >>= 1) {
				int ok = __shfl_xor(k, j), ov = __shfl_xor(v, j);
				bool up = (lane & kk) == 0, lower = (lane & j) == 0;
				if (lower == up ? ok < k : ok > k) {
					k = ok;
					v = ov;
				}
			}
		if (lane < e) put_sorted(rec[s + v], s + lane, c, nx, ny, d, perm, key, c1, c2, c3);
		return;
	}
	int *K = sk[w], *V = sv[w];
	if (e <= CS_MAX) {
		lds_bitonic(rec + s, e, K, V, lane);
		for (int i = lane; i < e; i += 64) put_sorted(rec[s + V[i]], s + i, c, nx, ny, d, perm, key, c1, c2, c3);
		return;
	}
	/* more than CS_MAX markers: four planes of cap ints in scratch (ids and
	 * slots twice), this cell's range at offset s in each */
	int *ka = scratch + s, *va = ka + cap, *kb = va + cap, *vb = kb + cap;
	for (int b = 0; b < e; b += CS_MAX) {
		const int len = min(CS_MAX, e - b);
		lds_bitonic(rec + s + b, len, K, V, lane);
		for (int i = lane; i < len; i += 64) {
			ka[b + i] = K[i];
			va[b + i] = b + V[i];
		}
		asm volatile("" ::: "memory");
	}
	for (int wd = CS_MAX; wd < e; wd <<= 1) {
		/* the previous pass's stores, by other lanes, before this pass's
		 * loads (agent scope also invalidates the vector L1) */
		__threadfence();
		for (int lo = 0; lo < e; lo += 2 * wd) {
			const int na = min(wd, e - lo), nb = max(0, min(wd, e - lo - wd));
			wave_merge(ka + lo, va + lo, na, ka + lo + na, va + lo + na, nb, kb + lo, vb + lo, lane);
		}
		int *t = ka; ka = kb; kb = t;
		t = va; va = vb; vb = t;
	}
	__threadfence();
	for (int i = lane; i < e; i += 64) put_sorted(rec[s + va[i]], s + i, c, nx, ny, d, perm, key, c1, c2, c3);
}

/* ------------------------------------------------------------------ */
/* Marker properties (ronurecalc per-marker part)                     */
/* ------------------------------------------------------------------ */

/* P at a point by prcalc1 (mark3mg.c:2917-3001): trilinear in cell
 * centres, cell indices from the (already clamped) search. */
__device__ double pr_interp(const double *pr, const double *gx, const double *gy, const double *gz, int nx,
			    int ny, int nz, double x, double y, double z, int m1, int m2, int m3)
{
	int a = m1, b = m2, c = m3;
	if (x > (gx[a] + gx[a + 1]) / 2.0) a += 1;
	if (a < 1) a = 1;
	if (a > nx - 2) a = nx - 2;
	if (y > (gy[b] + gy[b + 1]) / 2.0) b += 1;
	if (b < 1) b = 1;
	if (b > ny - 2) b = ny - 2;
	if (z > (gz[c] + gz[c + 1]) / 2.0) c += 1;
	if (c < 1) c = 1;
	if (c > nz - 2) c = nz - 2;
	double dx = (gx[a + 1] - gx[a - 1]) / 2.0, dy = (gy[b + 1] - gy[b - 1]) / 2.0, dz = (gz[c + 1] - gz[c - 1]) / 2.0;
	double p = 0;
	for (int m30 = c; m30 <= c + 1; m30++) {
		double ddz = ((gz[m30 - 1] + gz[m30]) / 2.0 - z) / dz;
		if (m30 == c) ddz = -ddz;
		ddz = 1.0 - ddz;
		for (int m10 = a; m10 <= a + 1; m10++) {
			double ddx = ((gx[m10 - 1] + gx[m10]) / 2.0 - x) / dx;
			if (m10 == a) ddx = -ddx;
			ddx = 1.0 - ddx;
			for (int m20 = b; m20 <= b + 1; m20++) {
				double ddy = ((gy[m20 - 1] + gy[m20]) / 2.0 - y) / dy;
				if (m20 == b) ddy = -ddy;
				ddy = 1.0 - ddy;
				long m40 = (long)m30 * nx * ny + (long)m10 * ny + m20;
				p += pr[m40] * (ddx * ddy * ddz);
			}
		}
	}
	return p;
}

/* viscalc branches A/B with the final limits (mark3mg.c:1266-1299,
 * 1626-1681); brittle and power-law terms are zero for these rocks. */
__device__ double visc_ab(const i3g_rock &r, double mtk, double mpb, double nubeg, double nuend)
{
	double rt = 8.314 * mtk, nunewt = 0;
	if (r.nu0 != 0) {
		if (r.dh == 0 && r.dv == 0) nunewt = 1.0 / r.nu0;
		else {
			double e1 = (r.dh + r.dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			if (r.dh < 0 && r.dv > 0) {
				e1 = r.dh * (mtk - r.dv);
				if (e1 < -150.0) e1 = -150.0;
			}
			if (r.dh < 0 && r.dv < 0) {
				e1 = (-r.dh) * (1.0 - (mtk - (-r.dv)) / (-r.dv)) / 8.314 / (-r.dv);
				if (e1 > 150.0) e1 = 150.0;
			}
			nunewt = 1.0 / (r.nu0 * exp(e1));
		}
	}
	double nueff = nunewt;
	if (nueff <= 0) nueff = 1.0 / r.n1;
	nueff = 1.0 / nueff;
	if (nueff < r.n0) nueff = r.n0;
	if (nueff > r.n1) nueff = r.n1;
	if (nueff < nubeg) nueff = nubeg;
	if (nueff > nuend) nueff = nuend;
	return nueff;
}

/* One marker's transfer record, written by k_marker_props in sorted order
 * and read whole by k_marker_gather: 80 bytes on one or two cache lines
 * per marker instead of one line per property array. */
struct __align__(16) MkRec {
	float x, y, z;
	int in; /* participates in the transfer */
	double nu, ro, cp, kt, ht, bb, tk;
	double pad;
};
static_assert(sizeof(MkRec) == 80 && sizeof(MkRec) % 16 == 0, "MkRec is copied in 16-byte words");

__device__ void epscalc(const float *st12, long N, const double *pr, const double *gx, const double *gy,
			const double *gz, int nx, int ny, int nz, double x, double y, double z, double *o);

/* ronurecalc's per-marker part (mark3mg.c:460-800): participation, with
 * rheology the retyping, viscalc (from the stored viscosity when there is
 * one and timestep > 0), dencalc1 with the phase-transition factor, the
 * water/air proportions at the water level (the marker's own y cell sets
 * the marker spacing, where the reference reads an index array shared
 * between threads) and the viscosity caps; the record for the gather. */
__device__ __forceinline__ void
marker_props_one(long t, const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
		 const float *__restrict__ tk, float *eta, signed char *type, float *e, float *rr, float *rn,
		 const int *c1, const int *c2, const int *c3, const double *pr, const float *st12, long N,
		 const double *gx, const double *gy, const double *gz, int nx, int ny, int nz, const i3g_rock *rocks,
		 int nrock, const i3g_mk_params &prm, const int *id, const float *mw, long mw_n, MkRec *R, int *bad)
{
#pragma clang fp contract(off)
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	float fx = x[t], fy = y[t], fz = z[t];
	int mt = type[t];
	/* marker participates (mark3mg.c:460-461) */
	int in = fx > 0 && fy > 0 && fz > 0 && fx < xs && fy < ys && fz < zs && mt < 50 && tk[t] > 0 && mt < nrock;
	R->x = fx; R->y = fy; R->z = fz; R->in = in;
	if (!in) return;
	int mm2 = mt;
	if (prm.rheology) {
		/* erosion, sedimentation, water/air (mark3mg.c:465-472) */
		if (mm2 > 1 && fy < prm.eroslev) mm2 = 0;
		if (mm2 < 2 && fy > prm.sedilev) mm2 = prm.sedtype;
		if (mm2 < 2 && fy > prm.waterlev) mm2 = 1;
		if (mm2 < 2 && fy < prm.waterlev) mm2 = 0;
		if (mm2 != mt) type[t] = (signed char)mm2;
		if (mm2 >= nrock || prm.densimod > 1) {
			atomicOr(bad, 1);
			R->in = 0;
			return;
		}
	} else if (mm2 < 2 || mm2 == 7 || mm2 == 8 || (mm2 >= 16 && mm2 <= 18) || mm2 > 20) {
		/* Newtonian subset: types with air/water, eclogite or melt branches
		 * (mark3mg.c:463-474, 530-641, 660-686) are refused */
		atomicOr(bad, 1);
		R->in = 0;
		return;
	}
	const i3g_rock r = rocks[mm2];
	double mpb = pr_interp(pr, gx, gy, gz, nx, ny, nz, (double)fx, (double)fy, (double)fz, c1[t], c2[t], c3[t]) * 1e-5;
	double mtk = (double)tk[t];
	double mnu = eta[t];
	VisPrm vp{prm.nubeg, prm.nuend, prm.timestep, prm.timesum, prm.convend};
	double eii = 0, sii = 0, dis = 0;
	float me = e[t], mrr = rr[t], mrn = rn[t];
	if (mnu <= 0 || prm.timestep <= 0) {
		if (prm.rheology) {
			double o[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
			if (st12) epscalc(st12, N, pr, gx, gy, gz, nx, ny, nz, (double)fx, (double)fy, (double)fz, o);
			eii = inv2_dev(o, 0);
			sii = inv2_dev(o, 7);
			dis = o[13];
		}
		mnu = viscalc_dev(r, mm2, mtk, mpb, eii, sii, 1.0, dis, me, mrr, mrn, vp);
		eta[t] = (float)mnu;
	}
	double mbb;
	double mro = dencalc_dev(r, prm.densimod, mtk, mpb, mbb);
	if (prm.rheology) mro *= phase_rokf_dev(mm2, mtk, mpb, prm.teclmin, prm.teclmax);
	double mcp = r.cp;
	double mkt = (r.kt + r.kf / (mtk + 77.0)) * exp(r.kp * mpb);
	double mht = r.ht;
	if (prm.rheology && mm2 > 20) {
		/* partially molten rock (mark3mg.c:660-686) */
		int mid = id[t];
		bool has = mw && mid >= 0 && mid < mw_n;
		double mwa = has ? (double)mw[mid] : 0.0, mex = has ? (double)mw[mw_n + mid] : 0.0, xm;
		meltpart0_dev(r, rocks[mm2 == 34 ? 11 : mm2 - 20], mm2, mtk, mpb, mwa, mex, prm.densimod, prm.teclmin,
			      prm.teclmax, prm.maxpmelt, mbb, mro, mcp, mkt, xm);
		if ((mm2 == 27 || mm2 == 28) && xm > 0.001) {
			/* basaltic melt viscosity (Schott and Schmeling 1995) */
			double nm = 1e+15 * exp(2.5 + pow((1.0 - xm) / xm, 0.48) * (1.0 - xm));
			if (nm < mnu) {
				mnu = nm;
				if (mnu < r.n0) mnu = r.n0;
				if (mnu > r.n1) mnu = r.n1;
				if (mnu < prm.nubeg) mnu = prm.nubeg;
				if (mnu > prm.nuend) mnu = prm.nuend;
				eta[t] = (float)mnu;
			}
		}
	}
	if (prm.rheology && prm.waterlev >= 0 && mm2 < 2) {
		int wy = c2[t];
		double dywa = (gy[wy + 1] - gy[wy]) / 2.0 / ((double)prm.mnumy);
		double lev = ((double)fy - (prm.waterlev - dywa)) / (2.0 * dywa);
		if (lev < 1.0 && lev > 0) {
			const i3g_rock ra = rocks[0], rw = rocks[1];
			double abb, wbb;
			double anu = viscalc_dev(ra, 0, mtk, mpb, eii, sii, 1.0, dis, me, mrr, mrn, vp);
			double aro = dencalc_dev(ra, prm.densimod, mtk, mpb, abb);
			double wnu = viscalc_dev(rw, 1, mtk, mpb, eii, sii, 1.0, dis, me, mrr, mrn, vp);
			double wro = dencalc_dev(rw, prm.densimod, mtk, mpb, wbb);
			mnu = wnu * lev + anu * (1.0 - lev);
			mro = wro * lev + aro * (1.0 - lev);
			mbb = wbb * lev + abb * (1.0 - lev);
			mcp = rw.cp * lev + ra.cp * (1.0 - lev);
			mkt = rw.kt * lev + ra.kt * (1.0 - lev);
			mht = rw.ht * lev + ra.ht * (1.0 - lev);
		}
	}
	if (prm.rheology) {
		e[t] = me;
		rr[t] = mrr;
		rn[t] = mrn;
	}
	/* viscosity caps of mark3mg.c:790-793 */
	if ((double)fy > prm.vdeep && mnu > prm.nudeep) mnu = prm.nudeep;
	if (fy < 5e+4 && fx > xs - 1e+4 && fz > 1.5e+5 && fz < 8.7e+5 && mnu > 1e+19) mnu = 1e+19;
	R->nu = mnu; R->ro = mro; R->cp = mcp; R->kt = mkt; R->ht = mht; R->bb = mbb; R->tk = mtk;
}

/* one thread per marker (marker_props_one); the block's records are
 * assembled in LDS and stored as one contiguous run with unit-stride
 * 16-byte stores (each thread writing its own 80-byte record directly
 * leaves partially written lines to the memory side) */
__global__ void __launch_bounds__(256)
k_marker_props(long n, const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
	       const float *__restrict__ tk, float *eta, signed char *type, float *e, float *rr, float *rn,
	       const int *c1, const int *c2, const int *c3, const double *pr, const float *st12, long N,
	       const double *gx, const double *gy, const double *gz, int nx, int ny, int nz, const i3g_rock *rocks,
	       int nrock, i3g_mk_params prm, const int *id, const float *mw, long mw_n, MkRec *rec, int *bad)
{
	__shared__ MkRec srec[256];
	long b0 = (long)blockIdx.x * blockDim.x, t = b0 + threadIdx.x;
	MkRec R;
	R.x = R.y = R.z = 0.f;
	R.in = 0;
	R.nu = R.ro = R.cp = R.kt = R.ht = R.bb = R.tk = R.pad = 0.0;
	if (t < n)
		marker_props_one(t, x, y, z, tk, eta, type, e, rr, rn, c1, c2, c3, pr, st12, N, gx, gy, gz, nx, ny, nz,
				 rocks, nrock, prm, id, mw, mw_n, &R, bad);
	srec[threadIdx.x] = R;
	__syncthreads();
	long cnt = n - b0 < (long)blockDim.x ? n - b0 : (long)blockDim.x;
	const uint4 *src = (const uint4 *)srec;
	uint4 *dst = (uint4 *)(rec + b0);
	int words = (int)(cnt * (long)(sizeof(MkRec) / 16));
	for (int w = threadIdx.x; w < words; w += blockDim.x) dst[w] = src[w];
}

/* ------------------------------------------------------------------ */
/* Marker -> node gather                                              */
/* ------------------------------------------------------------------ */

/* The markers of the (up to) eight cells around node (m1, m2, m3) in
 * increasing marker id: the cells' id-sorted runs merged.  Slot
 * q = 4*dc + 2*da + db is cell (m1-1+da, m2-1+db, m3-1+dc); a cell outside
 * the grid is an empty run.  The run heads' ids stay in registers and every
 * slot index is static, so one step issues a single id load (the advanced
 * run's next head), independent of the chosen marker's own loads. */
struct CellMerge {
	int ps[8], pe[8], hid[8];
	__device__ __forceinline__ void init(int nx, int ny, int nz, int m1, int m2, int m3, const int *start,
					     const int *cnt, const int *id)
	{
#pragma unroll
		for (int q = 0; q < 8; q++) {
			int a = m1 - 1 + ((q >> 1) & 1), b = m2 - 1 + (q & 1), c = m3 - 1 + (q >> 2);
			bool ok = a >= 0 && b >= 0 && c >= 0 && a <= nx - 2 && b <= ny - 2 && c <= nz - 2;
			long k = ok ? (long)c * nx * ny + (long)a * ny + b : 0;
			int s0 = ok ? start[k] : 0;
			ps[q] = s0;
			pe[q] = ok ? s0 + cnt[k] : 0;
		}
#pragma unroll
		for (int q = 0; q < 8; q++) hid[q] = ps[q] < pe[q] ? id[ps[q]] : 0x7fffffff;
	}
	/* next marker (sorted position mk) and its slot, and the load of that
	 * run's next head id into nid; false when every run is done.
	 * commit(slot, nid) after the marker's work applies the new head, so
	 * the id load overlaps the marker's own loads. */
	__device__ __forceinline__ bool pop(const int *id, int &mk, int &slot, int &nid)
	{
		int best = 0, bid = hid[0];
#pragma unroll
		for (int q = 1; q < 8; q++)
			if (hid[q] < bid) { bid = hid[q]; best = q; }
		if (bid == 0x7fffffff) return false;
		int p = ps[0], e = pe[0];
#pragma unroll
		for (int q = 1; q < 8; q++) {
			p = q == best ? ps[q] : p;
			e = q == best ? pe[q] : e;
		}
		nid = p + 1 < e ? id[p + 1] : 0x7fffffff;
		/* unconditional selects: every store keeps a static index, so the
		 * arrays stay in registers */
#pragma unroll
		for (int q = 0; q < 8; q++) ps[q] = q == best ? p + 1 : ps[q];
		mk = p;
		slot = best;
		return true;
	}
	__device__ __forceinline__ void commit(int slot, int nid)
	{
#pragma unroll
		for (int q = 0; q < 8; q++) hid[q] = q == slot ? nid : hid[q];
	}
};

/* widths of the two cells on either side of node index m along one axis
 * (zero where the cell is outside the grid: its slots are empty runs) */
__device__ __forceinline__ void side_widths(const double *g, int m, int n, double w[2])
{
	w[0] = m > 0 ? g[m] - g[m - 1] : 0.0;
	w[1] = m < n - 1 ? g[m + 1] - g[m] : 0.0;
}

/* Per node: merge the markers of the (up to) 8 cells around it by id and
 * accumulate the reference's weights (mark3mg.c:818-933) in that order;
 * then normalise (mark3mg.c:938-1002).  fld planes: 0 nu, 1 ro, 2 cp,
 * 3 kt, 4 ht, 5 et, 6 tk, 7 nxx, 8 nxy, 9 nxz, 10 nyz. */
__global__ void __launch_bounds__(256)
k_marker_gather(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz,
		const int *start, const int *cnt, const int *__restrict__ id, const MkRec *__restrict__ rec,
		double nukoef, double nubeg, double nuend, int viscmod, float *fld)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 >= nx) return;
	long N = (long)nx * ny * nz;
	long nd = (long)m3 * nx * ny + (long)m1 * ny + m2;
	CellMerge cm;
	cm.init(nx, ny, nz, m1, m2, m3, start, cnt, id);
	double wx[2], wy[2], wz[2];
	side_widths(gx, m1, nx, wx);
	side_widths(gy, m2, ny, wy);
	side_widths(gz, m3, nz, wz);
	double gxm = gx[m1], gym = gy[m2], gzm = gz[m3];
	float anu = 0, aro = 0, acp = 0, akt = 0, aht = 0, aet = 0, atk = 0, axx = 0, axy = 0, axz = 0, ayz = 0;
	double w1 = 0, w0 = 0, wxx = 0, wxy = 0, wxz = 0, wyz = 0;
	int mk, slot, nid;
	while (cm.pop(id, mk, slot, nid)) {
		/* every load of the step issued before one wait (the empty asm
		 * consumes them ahead of the branch on the participation flag) */
		const MkRec r = rec[mk];
		int in = r.in;
		float xm = r.x, ym = r.y, zm = r.z;
		double qnu = r.nu, qro = r.ro, qcp = r.cp, qkt = r.kt, qht = r.ht, qbb = r.bb, qtk = r.tk;
		asm volatile("" ::"v"(in), "v"(xm), "v"(ym), "v"(zm), "v"(qnu), "v"(qro), "v"(qcp), "v"(qkt), "v"(qht),
			     "v"(qbb), "v"(qtk));
		if (in) {
		int da = (slot >> 1) & 1, db = slot & 1, dc = slot >> 2;
		int a = m1 - 1 + da, b = m2 - 1 + db, c = m3 - 1 + dc;
		double dx = wx[da], dy = wy[db], dz = wz[dc];
		double ddz = 1.0 - fabs((gzm - (double)zm) / dz);
		double ddx = 1.0 - fabs((gxm - (double)xm) / dx);
		double ddy = 1.0 - fabs((gym - (double)ym) / dy);
		double swt = ddx * ddy * ddz;
		/* viscmod 0/1/2: arithmetic, geometric (log) or harmonic (inverse)
		 * averaging of the marker viscosity (mark3mg.c:848-900) */
		double mnu = viscmod == 1 ? log(qnu) : viscmod == 2 ? 1.0 / qnu : qnu;
		if (ddx > nukoef && ddy > nukoef && ddz > nukoef) {
			anu = (float)((double)anu + mnu * swt);
			w1 += swt;
		}
		if (m1 > a && m2 > b && m3 > c) {
			double s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (0.5 - fabs(ddy - 0.5)) * (0.5 - fabs(ddz - 0.5));
			axx = (float)((double)axx + mnu * s1);
			wxx += s1;
		}
		if (m3 == c && ddx > 0.5 && ddy > 0.5) {
			double s1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - fabs(ddz - 0.5));
			axy = (float)((double)axy + mnu * s1);
			wxy += s1;
		}
		if (m2 == b && ddx > 0.5 && ddz > 0.5) {
			double s1 = 8.0 * (ddx - 0.5) * (0.5 - fabs(ddy - 0.5)) * (ddz - 0.5);
			axz = (float)((double)axz + mnu * s1);
			wxz += s1;
		}
		if (m1 == a && ddy > 0.5 && ddz > 0.5) {
			double s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
			ayz = (float)((double)ayz + mnu * s1);
			wyz += s1;
		}
		if (swt > 0) {
			aro = (float)((double)aro + qro * swt);
			acp = (float)((double)acp + qcp * qro * swt);
			akt = (float)((double)akt + qkt * swt);
			aht = (float)((double)aht + qht * swt);
			aet = (float)((double)aet + qbb * swt);
			atk = (float)((double)atk + qtk * swt);
			w0 += swt;
		}
		}
		cm.commit(slot, nid);
	}
	float *f = fld;
	/* the mean back from the log / inverse (mark3mg.c:946-981) */
#define VMEAN(a, w) (viscmod == 1 ? (float)exp((double)(float)((double)(a) / (w)))                    \
		     : viscmod == 2 ? (float)(1.0 / (double)(float)((double)(a) / (w))) : (float)((double)(a) / (w)))
	if (w1 > 0.01) f[nd] = VMEAN(anu, w1);
	if (wxx > 0.01) f[7 * N + nd] = VMEAN(axx, wxx);
	if (wxy > 0.01) f[8 * N + nd] = VMEAN(axy, wxy);
	if (wxz > 0.01) f[9 * N + nd] = VMEAN(axz, wxz);
	if (wyz > 0.01) f[10 * N + nd] = VMEAN(ayz, wyz);
#undef VMEAN
	if (w0 > 0.01) {
		float rn = (float)((double)aro / w0);
		f[N + nd] = rn;
		f[2 * N + nd] = (float)((double)acp / w0 / (double)rn);
		f[3 * N + nd] = (float)((double)akt / w0);
		f[4 * N + nd] = (float)((double)aht / w0);
		f[5 * N + nd] = (float)((double)aet / w0);
		f[6 * N + nd] = (float)((double)atk / w0);
	}
	float nv = f[nd];
	if (nv < nubeg) nv = (float)nubeg;
	if (nv > nuend) nv = (float)nuend;
	f[nd] = nv;
}

/* ---- cell-segmented transfer ------------------------------------------- */

/* The markers of a cell are one contiguous id-ordered segment of the sorted
 * arrays.  k_cell_partials gives each cell eight lanes of a wavefront (lane
 * q = 4 ec + 2 ea + eb is the cell's corner node (a+ea, b+eb, c+ec)); the
 * eight lanes walk the segment together, each marker record loaded once for
 * the eight (one address per group), and each lane sums the reference's
 * weighted contributions (mark3mg.c:818-933) to its corner in FP64 in the
 * segment's id order.  k_node_sums adds, per node, the partials of its
 * (up to) eight cells in a fixed order and normalises as k_marker_gather.
 * Every record is read once (instead of by each of its eight nodes) and the
 * result does not depend on scheduling; it differs from the reference's
 * serial float accumulation (k_marker_gather) by that accumulation's own
 * rounding.  Partials: field F of corner q of cell s at
 * part[(F * 8 + q) * NC + s], NC cells in (c, a, b) order; a field a corner
 * never receives is not written. */
enum { PF_NU, PF_RO, PF_CP, PF_KT, PF_HT, PF_ET, PF_TK, PF_XX, PF_XY, PF_XZ, PF_YZ,
       PF_W1, PF_W0, PF_WXX, PF_WXY, PF_WXZ, PF_WYZ, PF_N };

/* which partial fields corner q receives: 0-6, w1, w0 always; xx at the
 * cell's far corner; xy with ec = 0; xz with eb = 0; yz with ea = 0 */
__device__ __forceinline__ bool pf_used(int F, int q)
{
	switch (F) {
	case PF_XX: case PF_WXX: return q == 7;
	case PF_XY: case PF_WXY: return (q >> 2) == 0;
	case PF_XZ: case PF_WXZ: return (q & 1) == 0;
	case PF_YZ: case PF_WYZ: return ((q >> 1) & 1) == 0;
	default: return true;
	}
}

__device__ __forceinline__ bool cell_of(long s, int nx, int ny, int nz, int *a, int *b, int *c)
{
	long hy = ny - 1, hx = nx - 1;
	if (s >= hx * hy * (nz - 1)) return false;
	*b = (int)(s % hy);
	long r = s / hy;
	*a = (int)(r % hx);
	*c = (int)(r / hx);
	return true;
}

/* One cell corner's share of a cell's markers: lane q = 4 ec + 2 ea + eb
 * of the cell's eight is corner node (a+ea, b+eb, c+ec). */
struct CornerGeom {
	double dx, dy, dz, gxm, gym, gzm;
	int q, ea, eb, ec;
	__device__ __forceinline__ void init(const double *gx, const double *gy, const double *gz, int a, int b, int c,
					     int lane)
	{
		q = lane;
		ea = (q >> 1) & 1;
		eb = q & 1;
		ec = q >> 2;
		dx = gx[a + 1] - gx[a];
		dy = gy[b + 1] - gy[b];
		dz = gz[c + 1] - gz[c];
		gxm = gx[a + ea];
		gym = gy[b + eb];
		gzm = gz[c + ec];
	}
};

/* the reference's weighted contributions of one marker record to corner
 * G.q (mark3mg.c:818-933), added to acc in FP64 */
__device__ __forceinline__ void corner_accum(const MkRec &r, const CornerGeom &G, double nukoef, int viscmod,
					     double *acc)
{
#pragma clang fp contract(off)
	if (!r.in) return;
	double ddz = 1.0 - fabs((G.gzm - (double)r.z) / G.dz);
	double ddx = 1.0 - fabs((G.gxm - (double)r.x) / G.dx);
	double ddy = 1.0 - fabs((G.gym - (double)r.y) / G.dy);
	double swt = ddx * ddy * ddz;
	double mnu = viscmod == 1 ? log(r.nu) : viscmod == 2 ? 1.0 / r.nu : r.nu;
	if (ddx > nukoef && ddy > nukoef && ddz > nukoef) {
		acc[PF_NU] += mnu * swt;
		acc[PF_W1] += swt;
	}
	if (G.q == 7) {
		double s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (0.5 - fabs(ddy - 0.5)) * (0.5 - fabs(ddz - 0.5));
		acc[PF_XX] += mnu * s1;
		acc[PF_WXX] += s1;
	}
	if (G.ec == 0 && ddx > 0.5 && ddy > 0.5) {
		double s1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - fabs(ddz - 0.5));
		acc[PF_XY] += mnu * s1;
		acc[PF_WXY] += s1;
	}
	if (G.eb == 0 && ddx > 0.5 && ddz > 0.5) {
		double s1 = 8.0 * (ddx - 0.5) * (0.5 - fabs(ddy - 0.5)) * (ddz - 0.5);
		acc[PF_XZ] += mnu * s1;
		acc[PF_WXZ] += s1;
	}
	if (G.ea == 0 && ddy > 0.5 && ddz > 0.5) {
		double s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
		acc[PF_YZ] += mnu * s1;
		acc[PF_WYZ] += s1;
	}
	if (swt > 0) {
		acc[PF_RO] += r.ro * swt;
		acc[PF_CP] += r.cp * r.ro * swt;
		acc[PF_KT] += r.kt * swt;
		acc[PF_HT] += r.ht * swt;
		acc[PF_ET] += r.bb * swt;
		acc[PF_TK] += r.tk * swt;
		acc[PF_W0] += swt;
	}
}

__device__ __forceinline__ void store_partials(const double *acc, int q, long NC, long s, double *part)
{
#pragma unroll
	for (int F = 0; F < PF_N; F++)
		if (pf_used(F, q)) part[((long)F * 8 + q) * NC + s] = acc[F];
}

/* the partials from the records k_marker_props staged (I3G_MK_STAGED=1) */
__global__ void __launch_bounds__(256)
k_cell_partials(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz, const int *start,
		const int *cnt, const MkRec *__restrict__ rec, double nukoef, int viscmod, double *part)
{
	long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long s = g >> 3;
	int a, b, c;
	if (!cell_of(s, nx, ny, nz, &a, &b, &c)) return;
	long NC = (long)(nx - 1) * (ny - 1) * (nz - 1);
	CornerGeom G;
	G.init(gx, gy, gz, a, b, c, (int)(g & 7));
	long key = (long)c * nx * ny + (long)a * ny + b;
	int k0 = start[key], k1 = k0 + cnt[key];
	double acc[PF_N];
#pragma unroll
	for (int F = 0; F < PF_N; F++) acc[F] = 0;
	for (int k = k0; k < k1; k++) corner_accum(rec[k], G, nukoef, viscmod, acc);
	store_partials(acc, G.q, NC, s, part);
}

/* The default transfer in one pass: the cell's eight lanes take its
 * id-ordered segment in chunks of eight markers; lane j computes marker
 * k + j's properties (marker_props_one) into the group's LDS slots, then
 * every lane adds the chunk's eight records to its corner in id order.
 * No record goes through memory, and the sums are those of
 * k_marker_props + k_cell_partials to the bit (same records, same order).
 * The lane's FP64 sums wait in LDS while the properties are computed.
 * The property code (FP64 rheology, pressure interpolation) needs more
 * than 256 registers a lane; bounded to two waves per SIMD it spills 88
 * bytes a lane and still runs faster than one wave without spills (134M
 * markers at 129^3: 12.8 against 17.4 ms; the staged two-pass transfer
 * 15.2 ms). */
#define CT_CELLS 32 /* cells per 256-thread block */
__global__ void __launch_bounds__(256, 2)
k_cell_transfer(int nx, int ny, int nz, const int *start, const int *cnt, const float *__restrict__ x,
		const float *__restrict__ y, const float *__restrict__ z, const float *__restrict__ tk, float *eta,
		signed char *type, float *e, float *rr, float *rn, const int *c1, const int *c2, const int *c3,
		const double *pr, const float *st12, long N, const double *gx, const double *gy, const double *gz,
		const i3g_rock *rocks, int nrock, i3g_mk_params prm, const int *id, const float *mw, long mw_n,
		int *bad, double *part)
{
	__shared__ MkRec srec[CT_CELLS][8];
	__shared__ double sacc[PF_N][256];
	const int grp = threadIdx.x >> 3, lane = threadIdx.x & 7;
	long s = (long)blockIdx.x * CT_CELLS + grp;
	int a, b, c;
	if (!cell_of(s, nx, ny, nz, &a, &b, &c)) return;
	long NC = (long)(nx - 1) * (ny - 1) * (nz - 1);
	long key = (long)c * nx * ny + (long)a * ny + b;
	int k0 = start[key], k1 = k0 + cnt[key];
	double acc[PF_N];
#pragma unroll
	for (int F = 0; F < PF_N; F++) sacc[F][threadIdx.x] = 0;
	MkRec *slot = srec[grp];
	for (int k = k0; k < k1; k += 8) {
		MkRec R;
		R.x = R.y = R.z = 0.f;
		R.in = 0;
		R.nu = R.ro = R.cp = R.kt = R.ht = R.bb = R.tk = R.pad = 0.0;
		if (k + lane < k1)
			marker_props_one(k + lane, x, y, z, tk, eta, type, e, rr, rn, c1, c2, c3, pr, st12, N, gx, gy, gz, nx,
					 ny, nz, rocks, nrock, prm, id, mw, mw_n, &R, bad);
		/* the group's eight lanes are in one wavefront, whose LDS
		 * operations complete in program order; the empty asm keeps the
		 * compiler from moving them across each other */
		slot[lane] = R;
		asm volatile("" ::: "memory");
		CornerGeom G;
		G.init(gx, gy, gz, a, b, c, lane);
#pragma unroll
		for (int F = 0; F < PF_N; F++) acc[F] = sacc[F][threadIdx.x];
		const int e8 = k1 - k < 8 ? k1 - k : 8;
		for (int j = 0; j < e8; j++) corner_accum(slot[j], G, prm.nukoef, prm.viscmod, acc);
#pragma unroll
		for (int F = 0; F < PF_N; F++) sacc[F][threadIdx.x] = acc[F];
		asm volatile("" ::: "memory");
	}
#pragma unroll
	for (int F = 0; F < PF_N; F++) acc[F] = sacc[F][threadIdx.x];
	store_partials(acc, lane, NC, s, part);
}

__global__ void __launch_bounds__(256)
k_node_sums(int nx, int ny, int nz, const double *__restrict__ part, double nubeg, double nuend, int viscmod,
	    float *fld)
{
#pragma clang fp contract(off)
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 >= nx) return;
	long N = (long)nx * ny * nz, NC = (long)(nx - 1) * (ny - 1) * (nz - 1);
	long nd = (long)m3 * nx * ny + (long)m1 * ny + m2;
	double acc[PF_N];
#pragma unroll
	for (int F = 0; F < PF_N; F++) acc[F] = 0;
#pragma unroll
	for (int q = 0; q < 8; q++) {
		/* slot q: cell (m1-1+da, m2-1+db, m3-1+dc), the node its corner 7-q */
		int da = (q >> 1) & 1, db = q & 1, dc = q >> 2;
		int a = m1 - 1 + da, b = m2 - 1 + db, c = m3 - 1 + dc;
		if (a < 0 || b < 0 || c < 0 || a > nx - 2 || b > ny - 2 || c > nz - 2) continue;
		long s = ((long)c * (nx - 1) + a) * (ny - 1) + b;
		int qc = 7 - q;
#pragma unroll
		for (int F = 0; F < PF_N; F++)
			if (pf_used(F, qc)) acc[F] += part[((long)F * 8 + qc) * NC + s];
	}
	float *f = fld;
#define VMEAN(a, w) (viscmod == 1 ? (float)exp((double)(float)((a) / (w)))                    \
		     : viscmod == 2 ? (float)(1.0 / (double)(float)((a) / (w))) : (float)((a) / (w)))
	if (acc[PF_W1] > 0.01) f[nd] = VMEAN(acc[PF_NU], acc[PF_W1]);
	if (acc[PF_WXX] > 0.01) f[7 * N + nd] = VMEAN(acc[PF_XX], acc[PF_WXX]);
	if (acc[PF_WXY] > 0.01) f[8 * N + nd] = VMEAN(acc[PF_XY], acc[PF_WXY]);
	if (acc[PF_WXZ] > 0.01) f[9 * N + nd] = VMEAN(acc[PF_XZ], acc[PF_WXZ]);
	if (acc[PF_WYZ] > 0.01) f[10 * N + nd] = VMEAN(acc[PF_YZ], acc[PF_WYZ]);
#undef VMEAN
	double w0 = acc[PF_W0];
	if (w0 > 0.01) {
		float rn = (float)(acc[PF_RO] / w0);
		f[N + nd] = rn;
		f[2 * N + nd] = (float)(acc[PF_CP] / w0 / (double)rn);
		f[3 * N + nd] = (float)(acc[PF_KT] / w0);
		f[4 * N + nd] = (float)(acc[PF_HT] / w0);
		f[5 * N + nd] = (float)(acc[PF_ET] / w0);
		f[6 * N + nd] = (float)(acc[PF_TK] / w0);
	}
	float nv = f[nd];
	if (nv < nubeg) nv = (float)nubeg;
	if (nv > nuend) nv = (float)nuend;
	f[nd] = nv;
}

/* tk = c (+ k * tk1[partner]) on temperature boundary nodes, in float as
 * tkrecalc (heat3mg.c:849-899); tk1 is the tk plane before the update */
__global__ void k_tkbc(long N, const int *__restrict__ p, const float *__restrict__ c, const float *__restrict__ k,
		       const float *__restrict__ tk1, float *__restrict__ tk)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= N) return;
	int q = p[t];
	if (q == -2) return;
	float v = c[t];
	if (q >= 0) v = v + k[t] * tk1[q];
	tk[t] = v;
}

/* ------------------------------------------------------------------ */
/* Marker temperature update after the heat solve                     */
/* ------------------------------------------------------------------ */

/* tkcalc1 (mark3mg.c:2844-2910): cell, spacings, node numbers and weights
 * of the 8 nodes around a point inside the grid, in the reference's node
 * order (z, then x, then y). */
__device__ __forceinline__ void tk_weights(const double *gx, const double *gy, const double *gz, int nx, int ny,
					   int nz, double x, double y, double z, int *cell, double *d, long *wn,
					   double *wi)
{
	int a0 = serch(gx, nx, x), b0 = serch(gy, ny, y), c0 = serch(gz, nz, z);
	cell[0] = a0; cell[1] = b0; cell[2] = c0;
	d[0] = gx[a0 + 1] - gx[a0];
	d[1] = gy[b0 + 1] - gy[b0];
	d[2] = gz[c0 + 1] - gz[c0];
	int q = 0;
	for (int c = c0; c <= c0 + 1; c++) {
		double ddz = 1.0 - fabs((gz[c] - z) / d[2]);
		for (int a = a0; a <= a0 + 1; a++) {
			double ddx = 1.0 - fabs((gx[a] - x) / d[0]);
			for (int b = b0; b <= b0 + 1; b++) {
				double ddy = 1.0 - fabs((gy[b] - y) / d[1]);
				wi[q] = ddx * ddy * ddz;
				wn[q] = (long)c * nx * ny + (long)a * ny + b;
				q++;
			}
		}
	}
}

/* Per marker (titerate, heat3mg.c:604-701): fluid markers (type 50..99)
 * take the new grid temperature; markers of type < 50 strictly inside the
 * grid are initialised when new, relaxed by the subgrid diffusion and given
 * the interpolated temperature change.  tin flags the markers whose
 * diffusion change enters the nodal average; the record carries the flag,
 * the position and the change for k_tk_gather. */
struct __align__(16) TkRec {
	float x, y, z;
	int in;
	double dtk, pad;
};

__global__ void k_tk_marker(long n, const float *__restrict__ x, const float *__restrict__ y,
			    const float *__restrict__ z, const signed char *__restrict__ type, float *markk,
			    const double *gx, const double *gy, const double *gz, int nx, int ny, int nz,
			    const float *__restrict__ tk, const float *__restrict__ tk2, const float *__restrict__ kt,
			    const float *__restrict__ ro, const float *__restrict__ cp, double heatdif, double dt,
			    double zdeep, double tdeep, TkRec *rec, uint8_t *tin)
{
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	tin[i] = 0;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	float fx = x[i], fy = y[i], fz = z[i];
	TkRec *R = rec + i;
	R->x = fx; R->y = fy; R->z = fz; R->in = 0;
	int mt = type[i];
	if (!(fx > 0 && fy > 0 && fz > 0 && (double)fx < xs && (double)fy < ys && (double)fz < zs)) return;
	if (mt >= 100) return;
	int cell[3];
	double d[3], wi[8];
	long wn[8];
	tk_weights(gx, gy, gz, nx, ny, nz, fx, fy, fz, cell, d, wn, wi);
	double e2 = 0, e3 = 0;
	for (int q = 0; q < 8; q++) {
		e2 += wi[q] * tk[wn[q]];
		e3 += wi[q] * tk2[wn[q]];
	}
	if (mt >= 50) {
		markk[i] = (float)e2;
		return;
	}
	float mk = markk[i];
	if (mk <= 0) {
		mk = (float)e3;
		if (fy > zdeep && mk < tdeep) mk = (float)tdeep;
	}
	if (heatdif != 0) {
		double mkt = 0, rocp = 0, mtk = tk[wn[0]];
		for (int q = 0; q < 8; q++) {
			mkt += kt[wn[q]] * wi[q];
			rocp += (ro[wn[q]] * cp[wn[q]]) * wi[q];
			mtk = fmax(mtk, (double)tk[wn[q]]);
		}
		if ((cell[0] <= 0 || cell[0] >= nx - 2 || cell[1] <= 0 || cell[1] >= ny - 2 || cell[2] <= 0 ||
		     cell[2] >= nz - 2) &&
		    ((double)mk + e2 - e3) > mtk)
			mk = (float)(mtk - e2 + e3);
		double dd = e3 - (double)mk;
		double d1 = -heatdif * mkt / rocp * 2.0 * (1.0 / d[0] / d[0] + 1.0 / d[1] / d[1] + 1.0 / d[2] / d[2]) * dt;
		if (d1 < -150.0) d1 = -150.0;
		d1 = dd * (1.0 - exp(d1));
		mk += (float)d1;
		R->dtk = d1;
		R->in = 1;
		tin[i] = 1;
	}
	mk += (float)(e2 - e3);
	markk[i] = mk;
}

/* Per node: average of the markers' diffusion changes, weights summed in
 * marker id order over the (up to) 8 cells around the node, as the
 * reference's serial scatter (heat3mg.c:687-699, 711-728). */
__global__ void __launch_bounds__(256)
k_tk_gather(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz, const int *start,
	    const int *cnt, const int *__restrict__ id, const TkRec *__restrict__ rec, double *sn)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 >= nx) return;
	long nd = (long)m3 * nx * ny + (long)m1 * ny + m2;
	CellMerge cm;
	cm.init(nx, ny, nz, m1, m2, m3, start, cnt, id);
	double wx[2], wy[2], wz[2];
	side_widths(gx, m1, nx, wx);
	side_widths(gy, m2, ny, wy);
	side_widths(gz, m3, nz, wz);
	double gxm = gx[m1], gym = gy[m2], gzm = gz[m3];
	double s0 = 0, s1 = 0;
	int mk, slot, nid;
	while (cm.pop(id, mk, slot, nid)) {
		const TkRec r = rec[mk];
		int in = r.in;
		float xm = r.x, ym = r.y, zm = r.z;
		double qtk = r.dtk;
		asm volatile("" ::"v"(in), "v"(xm), "v"(ym), "v"(zm), "v"(qtk));
		if (in) {
			double dx = wx[(slot >> 1) & 1], dy = wy[slot & 1], dz = wz[slot >> 2];
			double ddz = 1.0 - fabs((gzm - (double)zm) / dz);
			double ddx = 1.0 - fabs((gxm - (double)xm) / dx);
			double ddy = 1.0 - fabs((gym - (double)ym) / dy);
			double w = ddx * ddy * ddz;
			s1 += qtk * w;
			s0 += w;
		}
		cm.commit(slot, nid);
	}
	if (s0 != 0) s1 /= s0;
	sn[nd] = s1;
}

/* The same nodal average cell-segmented (see k_cell_partials): per cell
 * and corner the FP64 sums of dtk w and w over the cell's id-ordered
 * segment, then per node the eight cells' partials in a fixed order.  The
 * reference accumulates in double too (sol0/sol1), so only the order of
 * the FP64 additions differs from k_tk_gather. */
__global__ void __launch_bounds__(256)
k_tk_cell_partials(int nx, int ny, int nz, const double *gx, const double *gy, const double *gz, const int *start,
		   const int *cnt, const TkRec *__restrict__ rec, double *part)
{
#pragma clang fp contract(off)
	long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
	long s = g >> 3;
	int q = (int)(g & 7), a, b, c;
	if (!cell_of(s, nx, ny, nz, &a, &b, &c)) return;
	long NC = (long)(nx - 1) * (ny - 1) * (nz - 1);
	int ea = (q >> 1) & 1, eb = q & 1, ec = q >> 2;
	long key = (long)c * nx * ny + (long)a * ny + b;
	int k0 = start[key], k1 = k0 + cnt[key];
	double dx = gx[a + 1] - gx[a], dy = gy[b + 1] - gy[b], dz = gz[c + 1] - gz[c];
	double gxm = gx[a + ea], gym = gy[b + eb], gzm = gz[c + ec];
	double s1 = 0, s0 = 0;
	for (int k = k0; k < k1; k++) {
		const TkRec r = rec[k];
		if (!r.in) continue;
		double ddz = 1.0 - fabs((gzm - (double)r.z) / dz);
		double ddx = 1.0 - fabs((gxm - (double)r.x) / dx);
		double ddy = 1.0 - fabs((gym - (double)r.y) / dy);
		double w = ddx * ddy * ddz;
		s1 += r.dtk * w;
		s0 += w;
	}
	part[(long)q * NC + s] = s1;
	part[(long)(8 + q) * NC + s] = s0;
}

__global__ void __launch_bounds__(256)
k_tk_node_sums(int nx, int ny, int nz, const double *__restrict__ part, double *sn)
{
#pragma clang fp contract(off)
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	int m3 = blockIdx.y;
	int m1 = t / ny, m2 = t - m1 * ny;
	if (m1 >= nx) return;
	long NC = (long)(nx - 1) * (ny - 1) * (nz - 1);
	long nd = (long)m3 * nx * ny + (long)m1 * ny + m2;
	double s1 = 0, s0 = 0;
#pragma unroll
	for (int q = 0; q < 8; q++) {
		int da = (q >> 1) & 1, db = q & 1, dc = q >> 2;
		int a = m1 - 1 + da, b = m2 - 1 + db, c = m3 - 1 + dc;
		if (a < 0 || b < 0 || c < 0 || a > nx - 2 || b > ny - 2 || c > nz - 2) continue;
		long s = ((long)c * (nx - 1) + a) * (ny - 1) + b;
		s1 += part[(long)(7 - q) * NC + s];
		s0 += part[(long)(15 - q) * NC + s];
	}
	if (s0 != 0) s1 /= s0;
	sn[nd] = s1;
}

/* Per marker: antidiffusion, markk -= sum of the averaged nodal changes
 * (heat3mg.c:733-776). */
__global__ void k_tk_anti(long n, const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
			  const uint8_t *__restrict__ tin, float *markk, const double *gx, const double *gy,
			  const double *gz, int nx, int ny, int nz, const double *__restrict__ sn)
{
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || !tin[i]) return;
	int cell[3];
	double d[3], wi[8], a = 0;
	long wn[8];
	tk_weights(gx, gy, gz, nx, ny, nz, x[i], y[i], z[i], cell, d, wn, wi);
	for (int q = 0; q < 8; q++) a -= sn[wn[q]] * wi[q];
	markk[i] += (float)a;
}

/* ------------------------------------------------------------------ */
/* Strain rates / stresses at markers and movemark's property update  */
/* ------------------------------------------------------------------ */

/* One axis of an epscalc1 interpolation block (mark3mg.c:2575-2600,
 * 2646-2662): points at cell centres (kind 0: base cell + 1 past the
 * mid-line, in [1, n-2]), grid lines (kind 1: base in [1, n-3]) or
 * mid-lines (kind 2: base cell - 1 below the mid-line, in [0, n-3]);
 * w[0..1]: the weights of base and base+1 clamped to [0, 1]. */
__device__ __forceinline__ int eps_axis(const double *g, int n, int cell, double x, int kind, double *w)
{
	int b = cell;
	double d, p0, p1;
	if (kind == 0) {
		if (x > (g[b] + g[b + 1]) / 2.0) b += 1;
		if (b < 1) b = 1;
		if (b > n - 2) b = n - 2;
		d = (g[b + 1] - g[b - 1]) / 2.0;
		p0 = (g[b - 1] + g[b]) / 2.0;
		p1 = (g[b] + g[b + 1]) / 2.0;
	} else if (kind == 1) {
		if (b < 1) b = 1;
		if (b > n - 3) b = n - 3;
		d = g[b + 1] - g[b];
		p0 = g[b];
		p1 = g[b + 1];
	} else {
		if (x < (g[b] + g[b + 1]) / 2.0) b -= 1;
		if (b < 0) b = 0;
		if (b > n - 3) b = n - 3;
		d = (g[b + 2] - g[b]) / 2.0;
		p0 = (g[b] + g[b + 1]) / 2.0;
		p1 = (g[b + 1] + g[b + 2]) / 2.0;
	}
	double dd0 = 1.0 - (-((p0 - x) / d)), dd1 = 1.0 - (p1 - x) / d;
	w[0] = dd0 < 0 ? 0 : dd0 > 1.0 ? 1.0 : dd0;
	w[1] = dd1 < 0 ? 0 : dd1 > 1.0 ? 1.0 : dd1;
	return b;
}

/* epscalc1 (mark3mg.c:2535-2838) at a point: o[0..5] eps xx yy zz xy xz
 * yz, o[6] P, o[7..12] sigma, o[13] dissipation; normal components and P
 * from the eight surrounding cell centres, shear components from their
 * edge nodes; the reference's summation order. */
__device__ void epscalc(const float *st12, long N, const double *pr, const double *gx, const double *gy,
			const double *gz, int nx, int ny, int nz, double x, double y, double z, double *o)
{
	const int K[4][3] = {{0, 0, 0}, {1, 1, 2}, {1, 2, 1}, {2, 1, 1}};
	for (int q = 0; q < 14; q++) o[q] = 0;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	if (x < 0) x = 0;
	if (x > xs) x = xs;
	if (y < 0) y = 0;
	if (y > ys) y = ys;
	if (z < 0) z = 0;
	if (z > zs) z = zs;
	int m1 = serch(gx, nx, x), m2 = serch(gy, ny, y), m3 = serch(gz, nz, z);
	long sx = ny, sz = (long)nx * ny;
	for (int blk = 0; blk < 4; blk++) {
		double wx[2], wy[2], wz[2];
		int a = eps_axis(gx, nx, m1, x, K[blk][0], wx);
		int b = eps_axis(gy, ny, m2, y, K[blk][1], wy);
		int c = eps_axis(gz, nz, m3, z, K[blk][2], wz);
		for (int k = 0; k < 2; k++)
			for (int i = 0; i < 2; i++)
				for (int j = 0; j < 2; j++) {
					long u = (c + k) * sz + (long)(a + i) * sx + (b + j);
					double swt = wx[i] * wy[j] * wz[k];
					if (blk == 0) {
						float e0 = st12[u], e1 = st12[N + u], e2 = st12[2 * N + u];
						float s0 = st12[6 * N + u], s1 = st12[7 * N + u], s2 = st12[8 * N + u];
						o[0] += e0 * swt;
						o[1] += e1 * swt;
						o[2] += e2 * swt;
						o[7] += s0 * swt;
						o[8] += s1 * swt;
						o[9] += s2 * swt;
						o[13] += (s0 * e0 + s1 * e1 + s2 * e2) * swt;
						o[6] += pr[u] * swt;
					} else {
						float es = st12[(2 + blk) * N + u], ss = st12[(8 + blk) * N + u];
						o[2 + blk] += es * swt;
						o[9 + blk] += ss * swt;
						o[13] += 2.0 * ss * es * swt;
					}
				}
	}
}

/* markers that movemark moves (mark3mg.c:61): inside the grid, or outgrid
 * != 1 and type < 50 or >= 100 */
__device__ __forceinline__ bool mk_moves(float fx, float fy, float fz, int mt, double xs, double ys, double zs,
					 int outgrid)
{
	return (fx >= 0 && fy >= 0 && fz >= 0 && (double)fx <= xs && (double)fy <= ys && (double)fz <= zs) ||
	       (outgrid != 1 && (mt < 50 || mt >= 100));
}

__global__ void k_epscalc(long n, const float *__restrict__ x, const float *__restrict__ y,
			  const float *__restrict__ z, const float *st12, long N, const double *pr, const double *gx,
			  const double *gy, const double *gz, int nx, int ny, int nz, double *out)
{
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double o[14];
	epscalc(st12, N, pr, gx, gy, gz, nx, ny, nz, (double)x[t], (double)y[t], (double)z[t], o);
	for (int q = 0; q < 14; q++) out[t * 14 + q] = o[q];
}

/* movemark's marker properties ahead of the position update
 * (mark3mg.c:61-145): for the markers that move, epscalc1 at the marker
 * (strain-rate and stress invariants, dissipation, P); markers of type < 50
 * with T > 0 get the strain-rate correction of numerical diffusion
 * (markers away from the three outer lines that have a viscosity,
 * :97-105), eta = viscalc, the strain growth while yielding and the
 * healing.  rheology 0 refuses the types whose transfer branches it does
 * not take (as k_marker_props). */
__global__ void k_mk_props(long n, const float *__restrict__ x, const float *__restrict__ y,
			   const float *__restrict__ z, const float *__restrict__ tk, float *eta, float *e, float *rr,
			   float *rn, const signed char *__restrict__ type, const float *st12, long N, const double *pr,
			   const double *gx, const double *gy, const double *gz, int nx, int ny, int nz,
			   const i3g_rock *rocks, int nrock, i3g_mk_params prm, double healrate, int *bad)
{
#pragma clang fp contract(off)
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	float fx = x[t], fy = y[t], fz = z[t];
	int mt = type[t];
	if (!mk_moves(fx, fy, fz, mt, gx[nx - 1], gy[ny - 1], gz[nz - 1], prm.outgrid)) return;
	double mtk = (double)tk[t];
	if (!(mt < 50 && mtk > 0)) return;
	if (mt >= nrock || (!prm.rheology && (mt < 2 || mt == 7 || mt == 8 || (mt >= 16 && mt <= 18) || mt > 20))) {
		atomicOr(bad, 1);
		return;
	}
	double o[14];
	epscalc(st12, N, pr, gx, gy, gz, nx, ny, nz, (double)fx, (double)fy, (double)fz, o);
	double eii = inv2_dev(o, 0), sii = inv2_dev(o, 7), dis = o[13];
	double mpb = o[6] * 1e-5;
	float mv = eta[t];
	double rii = 1.0;
	if (mv > 0 && fx > gx[3] && fx < gx[nx - 4] && fy > gy[3] && fy < gy[ny - 4] && fz > gz[3] && fz < gz[nz - 4])
		rii *= pow(0.5 * sii / eii / mv, 0.87);
	float me = e[t], mrr = rr[t], mrn = rn[t];
	VisPrm vp{prm.nubeg, prm.nuend, prm.timestep, prm.timesum, prm.convend};
	eta[t] = (float)viscalc_dev(rocks[mt], mt, mtk, mpb, eii, sii, rii, dis, me, mrr, mrn, vp);
	if (me > 0) me += prm.timestep * eii;
	if (me > 0) {
		me -= prm.timestep * healrate;
		if (me < 0) me = 0;
	} else {
		me += prm.timestep * healrate;
		if (me > 0) me = 0;
	}
	e[t] = me;
	rr[t] = mrr;
	rn[t] = mrn;
}

/* ------------------------------------------------------------------ */
/* RK4 advection                                                      */
/* ------------------------------------------------------------------ */

/* zdeep / vdeep / tdeep of movemark's depth rules (mark3mg.c:254-276):
 * below zdeep a sinking marker becomes type 10 (at least tdeep when it is
 * already 10), any other below zdeep type 9; a marker rising through
 * vdeep or zdeep becomes type 10 with T at least tdeep. */
struct DepthRules {
	double zdeep, vdeep, tdeep;
};

/* movemark's motion of one marker (mark3mg.c:61, 146-284) */
struct MoveArgs {
	i3g_move_params mp;
	const double *pr; /* P (level-0 nodes) for the fluid velocity, or null */
	const float *md;  /* marker density by id, or null */
	const int *id;
};

__global__ void k_rk4(long n, float *x, float *y, float *z, float *tk, float *eta, signed char *type,
		      const float *vx, const float *vy, const float *vz, const double *gx, const double *gy,
		      const double *gz, int nx, int ny, int nz, double dt, DepthRules D, MoveArgs A)
{
#pragma clang fp contract(off)
	extern __shared__ double sg[];
	stage_lines(sg, gx, gy, gz, nx, ny, nz);
	gx = sg; gy = sg + nx; gz = sg + nx + ny;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	float fx = x[t], fy = y[t], fz = z[t];
	int mt = type[t];
	/* moved markers (mark3mg.c:61) */
	if (!((fx >= 0 && fy >= 0 && fz >= 0 && (double)fx <= xs && (double)fy <= ys && (double)fz <= zs) ||
	      (A.mp.outgrid != 1 && (mt < 50 || mt >= 100))))
		return;
	double px = fx, py = fy, pz = fz;
	int cell[3];
	double v1[3], v2[3], v3[3], v4[3], v[3], vw[3] = {0, 0, 0};
	vxyz(vx, vy, vz, gx, gy, gz, nx, ny, nz, px, py, pz, v1, cell, false);
	if (mt >= 50 && mt < 100) {
		/* water velocity of fluid markers (mark3mg.c:146-207) */
		float ftk = tk[t];
		vw[1] = A.mp.vyfluid;
		if (A.md && A.md[A.id[t]] > 1100.0f) vw[1] = A.mp.vymelt;
		if (A.mp.vyfluid > 0 && (ftk == 0 || ftk > 298.0f)) {
			int m1 = cell[0], m2 = cell[1], m3 = cell[2];
			if ((double)fx > (gx[m1] + gx[m1 + 1]) / 2.0) m1 += 1;
			if (m1 < 1) m1 = 1;
			if (m1 > nx - 2) m1 = nx - 2;
			if ((double)fy > (gy[m2] + gy[m2 + 1]) / 2.0) m2 += 1;
			if (m2 < 1) m2 = 1;
			if (m2 > ny - 2) m2 = ny - 2;
			if ((double)fz > (gz[m3] + gz[m3 + 1]) / 2.0) m3 += 1;
			if (m3 < 1) m3 = 1;
			if (m3 > nz - 2) m3 = nz - 2;
			double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
			double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
			double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
			/* the reference scales all three offsets by the x factor */
			double ax = xkf * ((double)fx - (gx[m1 - 1] + gx[m1]) / 2.0);
			double ay = xkf * ((double)fy - (gy[m2 - 1] + gy[m2]) / 2.0);
			double az = xkf * ((double)fz - (gz[m3 - 1] + gz[m3]) / 2.0);
			long N = (long)nx * ny * nz, sy = ny, sxy = (long)nx * ny;
			long p[8];
			p[0] = (long)m3 * sxy + (long)m1 * sy + m2;
			p[1] = p[0] + 1;
			p[2] = p[0] + sy;
			p[3] = p[2] + 1;
			p[4] = p[0] + sxy;
			p[5] = p[4] + 1;
			p[6] = p[5] + sy; /* one y line above p[4] + sy, as the reference */
			p[7] = p[6] + 1;
			double pv[8];
			for (int q = 0; q < 8; q++) pv[q] = (A.pr && p[q] < N) ? A.pr[p[q]] : 0.0;
			double d[3];
			d[0] = xkf * ((1.0 - ay) * (1.0 - az) * (pv[2] - pv[0]) + ay * (1.0 - az) * (pv[3] - pv[1]) +
				      (1.0 - ay) * az * (pv[6] - pv[4]) + ay * az * (pv[7] - pv[5]));
			d[1] = ykf * ((1.0 - ax) * (1.0 - az) * (pv[1] - pv[0]) + ax * (1.0 - az) * (pv[3] - pv[2]) +
				      (1.0 - ax) * az * (pv[5] - pv[4]) + ax * az * (pv[7] - pv[6]));
			d[2] = zkf * ((1.0 - ax) * (1.0 - ay) * (pv[4] - pv[0]) + ax * (1.0 - ay) * (pv[6] - pv[2]) +
				      (1.0 - ax) * ay * (pv[5] - pv[1]) + ax * ay * (pv[7] - pv[3]));
			double k[3];
			for (int q = 0; q < 3; q++) {
				k[q] = (1000.0 * A.mp.g[q] - d[q]) / (2300.0 * 9.81);
				if (k[q] > 2.0) k[q] = 2.0;
				if (k[q] < -2.0) k[q] = -2.0;
			}
			vw[0] = vw[1] * k[0];
			vw[2] = vw[1] * k[2];
			vw[1] *= k[1];
		} else {
			vw[0] = 0;
			vw[1] = -fabs(vw[1]);
			vw[2] = 0;
		}
	}
	if (A.mp.markmod == 1) {
		for (int q = 0; q < 3; q++) v[q] = v1[q] + vw[q];
	} else {
		/* stages 2-4 search their cell from the previous stage's (a step of
		 * a fraction of a cell: one or two comparisons instead of a
		 * bisection) */
		for (int q = 0; q < 3; q++) v1[q] += vw[q];
		vxyz(vx, vy, vz, gx, gy, gz, nx, ny, nz, px + v1[0] * dt / 2.0, py + v1[1] * dt / 2.0, pz + v1[2] * dt / 2.0,
		     v2, cell, true);
		for (int q = 0; q < 3; q++) v2[q] += vw[q];
		vxyz(vx, vy, vz, gx, gy, gz, nx, ny, nz, px + v2[0] * dt / 2.0, py + v2[1] * dt / 2.0, pz + v2[2] * dt / 2.0,
		     v3, cell, true);
		for (int q = 0; q < 3; q++) v3[q] += vw[q];
		vxyz(vx, vy, vz, gx, gy, gz, nx, ny, nz, px + v3[0] * dt, py + v3[1] * dt, pz + v3[2] * dt, v4, cell, true);
		for (int q = 0; q < 3; q++) v4[q] += vw[q];
		for (int q = 0; q < 3; q++) v[q] = (v1[q] + 2.0 * v2[q] + 2.0 * v3[q] + v4[q]) / 6.0;
	}
	/* motion normal to the faces only (mark3mg.c:245-251) */
	if (A.mp.outgrid == 2) {
		if (fx < 0 || (double)fx > xs) v[1] = v[2] = 0;
		if (fy < 0 || (double)fy > ys) v[0] = v[2] = 0;
		if (fz < 0 || (double)fz > zs) v[0] = v[1] = 0;
	}
	if (mt < 100) {
		float ftk = tk[t];
		double yold = fy;
		if ((double)fy > D.zdeep && v[1] < 0 && (double)ftk < D.tdeep && mt == 10) ftk = (float)D.tdeep;
		if ((double)fy > D.zdeep && v[1] < 0 && mt != 10) mt = 10;
		if ((double)fy > D.zdeep && mt != 10 && mt != 9) mt = 9;
		fx = (float)((double)fx + (dt * v[0] - 0.0));
		fy = (float)((double)fy + dt * v[1]);
		fz = (float)((double)fz + dt * v[2]);
		x[t] = fx; y[t] = fy; z[t] = fz;
		if ((yold > D.vdeep && (double)fy < D.vdeep) || (yold > D.zdeep && (double)fy < D.zdeep)) {
			if ((double)ftk < D.tdeep) ftk = (float)D.tdeep;
			mt = 10;
		}
		tk[t] = ftk;
		type[t] = (signed char)mt;
		if (fx < 0 || fy < 0 || fz < 0 || (double)fx > xs || (double)fy > ys || (double)fz > zs) {
			tk[t] = 0;
			eta[t] = 0;
		}
	} else {
		/* immobile markers only move; their reset stays with the caller */
		x[t] = (float)((double)fx + (dt * v[0] - 0.0));
		y[t] = (float)((double)fy + dt * v[1]);
		z[t] = (float)((double)fz + dt * v[2]);
	}
}

/* ------------------------------------------------------------------ */
/* C ABI                                                              */
/* ------------------------------------------------------------------ */

template <typename T> static T *mk_alloc(size_t n, int *err)
{
	void *p = nullptr;
	if (n == 0) n = 1;
	if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { *err = I3G_ENOMEM; return nullptr; }
	return (T *)p;
}

static inline dim3 g1(long n) { return dim3((unsigned)((n + 255) / 256)); }
static inline unsigned lines_bytes(const i3g_markers *m) { return (unsigned)((m->nx + m->ny + m->nz) * sizeof(double)); }

extern "C" int i3g_markers_create(long nx, long ny, long nz, const double *gx, const double *gy, const double *gz,
				  long capacity, i3g_markers **out)
{
	if (!out || !gx || !gy || !gz || nx < 3 || ny < 3 || nz < 3 || capacity < 1 || capacity > 2000000000L)
		return I3G_EINVAL;
	if (nx + ny + nz > 6144) return I3G_ECAP; /* grid lines staged in 48 KB of LDS */
	i3g_markers *m = new i3g_markers();
	m->nx = (int)nx; m->ny = (int)ny; m->nz = (int)nz;
	m->N = nx * ny * nz;
	m->cap = capacity;
	if (const char *ev = getenv("I3G_MK_STAGED")) m->staged = ev[0] == '1';
	int err = 0;
	if (hipStreamCreate(&m->s) != hipSuccess) { delete m; return I3G_EDEVICE; }
	if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess) {
		i3g_markers_destroy(m);
		return I3G_EDEVICE;
	}
	m->gx = mk_alloc<double>(nx, &err); m->gy = mk_alloc<double>(ny, &err); m->gz = mk_alloc<double>(nz, &err);
	long c = capacity;
	m->x = mk_alloc<float>(c, &err); m->y = mk_alloc<float>(c, &err); m->z = mk_alloc<float>(c, &err);
	m->tk = mk_alloc<float>(c, &err); m->eta = mk_alloc<float>(c, &err);
	m->e = mk_alloc<float>(c, &err); m->e2 = mk_alloc<float>(c, &err);
	m->rr = mk_alloc<float>(c, &err); m->rr2 = mk_alloc<float>(c, &err);
	m->rn = mk_alloc<float>(c, &err); m->rn2 = mk_alloc<float>(c, &err);
	m->type = mk_alloc<signed char>(c, &err); m->id = mk_alloc<int>(c, &err);
	m->x2 = mk_alloc<float>(c, &err); m->y2 = mk_alloc<float>(c, &err); m->z2 = mk_alloc<float>(c, &err);
	m->tk2 = mk_alloc<float>(c, &err); m->eta2 = mk_alloc<float>(c, &err);
	m->type2 = mk_alloc<signed char>(c, &err); m->id2 = mk_alloc<int>(c, &err);
	m->key = mk_alloc<int>(c, &err); m->c1 = mk_alloc<int>(c, &err); m->c2 = mk_alloc<int>(c, &err);
	m->c3 = mk_alloc<int>(c, &err); m->perm = mk_alloc<int>(c, &err);
	m->cnt = mk_alloc<int>(m->N, &err); m->start = mk_alloc<int>(m->N + 1, &err); m->cur = mk_alloc<int>(m->N, &err);
	m->bsum = mk_alloc<int>(m->N / SCAN_B + 2, &err);
	m->rocks = mk_alloc<i3g_rock>(MAXROCK, &err);
	m->bad = mk_alloc<int>(1, &err);
	if (err) { i3g_markers_destroy(m); return err; }
	if (hipMemcpy(m->gx, gx, nx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(m->gy, gy, ny * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(m->gz, gz, nz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
		i3g_markers_destroy(m);
		return I3G_EDEVICE;
	}
	*out = m;
	return I3G_OK;
}

extern "C" int i3g_markers_destroy(i3g_markers *m)
{
	if (!m) return I3G_OK;
	hipDeviceSynchronize();
	void *ps[] = {m->gx, m->gy, m->gz, m->x, m->y, m->z, m->tk, m->eta, m->type, m->id, m->x2, m->y2, m->z2,
		      m->tk2, m->eta2, m->type2, m->id2, m->key, m->c1, m->c2, m->c3, m->perm, m->cnt, m->start,
		      m->cur, m->bsum, m->prec, m->pr, m->fld,
		      m->vel, m->rocks, m->bad, m->tbc_p, m->tbc_c, m->tbc_k, m->tk1, m->tnod, m->trec, m->sn, m->tin,
		      m->e, m->e2, m->st12, m->rr, m->rr2, m->rn, m->rn2, m->md, m->mw, m->part};
	for (void *p : ps)
		if (p) hipFree(p);
	if (m->ev0) hipEventDestroy(m->ev0);
	if (m->ev1) hipEventDestroy(m->ev1);
	if (m->s) hipStreamDestroy(m->s);
	delete m;
	return I3G_OK;
}

extern "C" long i3g_markers_count(const i3g_markers *m) { return m ? m->n : I3G_EINVAL; }

extern "C" double i3g_markers_last_ms(const i3g_markers *m) { return m ? (double)m->last_ms : -1.0; }

extern "C" int i3g_markers_upload(i3g_markers *m, long n, const float *x, const float *y, const float *z,
				  const float *tk, const float *eta, const signed char *type, const int *id)
{
	if (!m || n < 0 || n > m->cap || !x || !y || !z || !tk || !eta || !type) return I3G_EINVAL;
	m->n = n;
	m->sorted = 0;
	MCHK(hipMemcpyAsync(m->x, x, n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->y, y, n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->z, z, n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->tk, tk, n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->eta, eta, n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->type, type, n * sizeof(signed char), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemsetAsync(m->e, 0, (n ? n : 1) * sizeof(float), m->s));
	MCHK(hipMemsetAsync(m->rr, 0, (n ? n : 1) * sizeof(float), m->s));
	MCHK(hipMemsetAsync(m->rn, 0, (n ? n : 1) * sizeof(float), m->s));
	if (id) {
		MCHK(hipMemcpyAsync(m->id, id, n * sizeof(int), hipMemcpyHostToDevice, m->s));
	} else {
		std::vector<int> ids(n);
		for (long i = 0; i < n; i++) ids[i] = (int)i;
		MCHK(hipMemcpyAsync(m->id, ids.data(), n * sizeof(int), hipMemcpyHostToDevice, m->s));
		MCHK(hipStreamSynchronize(m->s));
	}
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_download(i3g_markers *m, float *x, float *y, float *z, float *tk, float *eta,
				    signed char *type, int *id)
{
	if (!m) return I3G_EINVAL;
	long n = m->n;
	if (x) MCHK(hipMemcpyAsync(x, m->x, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	if (y) MCHK(hipMemcpyAsync(y, m->y, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	if (z) MCHK(hipMemcpyAsync(z, m->z, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	if (tk) MCHK(hipMemcpyAsync(tk, m->tk, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	if (eta) MCHK(hipMemcpyAsync(eta, m->eta, n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	if (type) MCHK(hipMemcpyAsync(type, m->type, n * sizeof(signed char), hipMemcpyDeviceToHost, m->s));
	if (id) MCHK(hipMemcpyAsync(id, m->id, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

static int cell_index(i3g_markers *m, int with_count)
{
	if (with_count) MCHK(hipMemsetAsync(m->cnt, 0, m->N * sizeof(int), m->s));
	if (m->n)
		hipLaunchKernelGGL(k_cell_index, g1(m->n), dim3(256), lines_bytes(m), m->s, m->n, (const float *)m->x,
				   (const float *)m->y, (const float *)m->z, (const double *)m->gx,
				   (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz, m->key, m->c1,
				   m->c2, m->c3, with_count ? m->cnt : (int *)nullptr);
	MCHK(hipGetLastError());
	return I3G_OK;
}

extern "C" int i3g_markers_cell_index(i3g_markers *m, int *m1, int *m2, int *m3, int *key)
{
	if (!m) return I3G_EINVAL;
	mark_begin(m);
	int rc = cell_index(m, 0);
	if (rc) return rc;
	mark_end(m);
	long n = m->n;
	if (m1) MCHK(hipMemcpyAsync(m1, m->c1, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
	if (m2) MCHK(hipMemcpyAsync(m2, m->c2, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
	if (m3) MCHK(hipMemcpyAsync(m3, m->c3, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
	if (key) MCHK(hipMemcpyAsync(key, m->key, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
	MCHK(hipStreamSynchronize(m->s));
	mark_read(m);
	return I3G_OK;
}

/* the sorted arrays are in the second buffers: swap */
static void swap_sorted(i3g_markers *m)
{
	std::swap(m->x, m->x2); std::swap(m->y, m->y2); std::swap(m->z, m->z2);
	std::swap(m->tk, m->tk2); std::swap(m->eta, m->eta2); std::swap(m->e, m->e2);
	std::swap(m->rr, m->rr2); std::swap(m->rn, m->rn2);
	std::swap(m->type, m->type2); std::swap(m->id, m->id2);
}

extern "C" int i3g_markers_sort(i3g_markers *m, long *order)
{
	if (!m) return I3G_EINVAL;
	long n = m->n, N = m->N;
	mark_begin(m);
	int rc = cell_index(m, 1);
	if (rc) return rc;
	long nb = (N + SCAN_B - 1) / SCAN_B;
	hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(SCAN_B), 0, m->s, (const int *)m->cnt, m->start, N, m->bsum);
	hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(SCAN_B), 0, m->s, m->bsum, nb);
	hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_B), 0, m->s, m->start, N, (const int *)m->bsum, 0L);
	MCHK(hipMemsetAsync(m->cur, 0, N * sizeof(int), m->s));
	if (n) {
		/* the records are staged in the node transfer's record buffer
		 * (rewritten by every transfer after a sort) */
		int err = 0;
		if (!m->prec) m->prec = mk_alloc<MkRec>(m->cap, &err);
		if (err) return err;
		static_assert(sizeof(SortRec) <= sizeof(MkRec), "sort records fit the transfer records");
		SortRec *rec = (SortRec *)m->prec;
		MkArrays src{m->x, m->y, m->z, m->tk, m->eta, m->e, m->rr, m->rn, m->type, m->id};
		MkArrays dst{m->x2, m->y2, m->z2, m->tk2, m->eta2, m->e2, m->rr2, m->rn2, m->type2, m->id2};
		hipLaunchKernelGGL(k_fill, g1(n), dim3(256), 0, m->s, n, (const int *)m->key, (const int *)m->start, m->cur,
				   src, rec);
		/* the scratch of cells beyond CS_MAX markers (four planes of cap
		 * ints, indexed by sorted position): the record buffer's tail past
		 * cap SortRecs */
		static_assert(sizeof(MkRec) - sizeof(SortRec) >= 4 * sizeof(int), "scratch fits behind the sort records");
		int *scratch = (int *)(rec + m->cap);
		hipLaunchKernelGGL(k_cellsort, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, m->s, N, m->cap, m->nx, m->ny,
				   (const int *)m->start, (const int *)m->cnt, (const SortRec *)rec, dst, m->perm, m->key,
				   m->c1, m->c2, m->c3, scratch);
		swap_sorted(m);
	}
	MCHK(hipGetLastError());
	mark_end(m);
	if (order && n) {
		std::vector<int> p(n);
		MCHK(hipMemcpyAsync(p.data(), m->perm, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
		MCHK(hipStreamSynchronize(m->s));
		for (long i = 0; i < n; i++) order[i] = p[i];
	}
	MCHK(hipStreamSynchronize(m->s));
	mark_read(m);
	m->sorted = 1;
	return I3G_OK;
}

/* the per-cell partials of the segmented node averages, allocated on first
 * use (PF_N x 8 doubles per cell; the temperature average uses 16) */
static int ensure_part(i3g_markers *m)
{
	if (m->order != 1 || m->part) return I3G_OK;
	int err = 0;
	m->part = mk_alloc<double>((size_t)PF_N * 8 * (m->nx - 1) * (m->ny - 1) * (m->nz - 1), &err);
	return err;
}

extern "C" int i3g_markers_set_order(i3g_markers *m, int order)
{
	if (!m || order < 0 || order > 1) return I3G_EINVAL;
	m->order = order;
	return I3G_OK;
}

extern "C" int i3g_markers_to_nodes(i3g_markers *m, const i3g_rock *rocks, int nrock, const i3g_mk_params *prm,
				    const double *pr, float *nu, float *ro, float *cp, float *kt, float *ht, float *et,
				    float *tk, float *nxx, float *nxy, float *nxz, float *nyz)
{
	if (!m || !rocks || !prm || !pr || nrock < 1 || nrock > MAXROCK) return I3G_EINVAL;
	if (prm->viscmod < 0 || prm->viscmod > 2 || prm->densimod < 0 || prm->densimod > 1) return I3G_EINVAL;
	if (prm->rheology && (prm->mnumy < 1 || prm->sedtype < 0 || prm->sedtype >= nrock)) return I3G_EINVAL;
	if (!m->sorted) return I3G_EINVAL;
	float *host[11] = {nu, ro, cp, kt, ht, et, tk, nxx, nxy, nxz, nyz};
	for (int q = 0; q < 11; q++)
		if (!host[q]) return I3G_EINVAL;
	int err = 0;
	long n = m->n, N = m->N;
	const bool fused = m->order == 1 && !m->staged;
	if (!fused && !m->prec) m->prec = mk_alloc<MkRec>(m->cap, &err);
	if (!m->fld) m->fld = mk_alloc<float>(11 * N, &err);
	if (err) return err;
	if (!m->pr) {
		m->pr = mk_alloc<double>(N, &err);
		if (err) return err;
	}
	if ((err = ensure_part(m))) return err;
	MCHK(hipMemcpyAsync(m->rocks, rocks, nrock * sizeof(i3g_rock), hipMemcpyDefault, m->s));
	MCHK(hipMemsetAsync(m->bad, 0, sizeof(int), m->s));
	MCHK(hipMemcpyAsync(m->pr, pr, N * sizeof(double), hipMemcpyDefault, m->s));
	for (int q = 0; q < 11; q++)
		MCHK(hipMemcpyAsync(m->fld + q * N, host[q], N * sizeof(float), hipMemcpyDefault, m->s));
	mark_begin(m);
	if (n && !fused)
		hipLaunchKernelGGL(k_marker_props, g1(n), dim3(256), 0, m->s, n, (const float *)m->x, (const float *)m->y,
				   (const float *)m->z, (const float *)m->tk, m->eta, m->type, m->e, m->rr, m->rn,
				   (const int *)m->c1, (const int *)m->c2, (const int *)m->c3, (const double *)m->pr,
				   (const float *)(m->has_st ? m->st12 : nullptr), N, (const double *)m->gx,
				   (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz, (const i3g_rock *)m->rocks,
				   nrock, *prm, (const int *)m->id, (const float *)m->mw, m->mw_n, m->prec, m->bad);
	dim3 gg((unsigned)(((long)m->nx * m->ny + 255) / 256), (unsigned)m->nz);
	if (m->order == 1) {
		long NC = (long)(m->nx - 1) * (m->ny - 1) * (m->nz - 1);
		if (fused)
			hipLaunchKernelGGL(k_cell_transfer, dim3((unsigned)((NC + CT_CELLS - 1) / CT_CELLS)), dim3(256), 0, m->s,
					   m->nx, m->ny, m->nz, (const int *)m->start, (const int *)m->cnt, (const float *)m->x,
					   (const float *)m->y, (const float *)m->z, (const float *)m->tk, m->eta, m->type, m->e,
					   m->rr, m->rn, (const int *)m->c1, (const int *)m->c2, (const int *)m->c3,
					   (const double *)m->pr, (const float *)(m->has_st ? m->st12 : nullptr), N,
					   (const double *)m->gx, (const double *)m->gy, (const double *)m->gz,
					   (const i3g_rock *)m->rocks, nrock, *prm, (const int *)m->id, (const float *)m->mw,
					   m->mw_n, m->bad, m->part);
		else
			hipLaunchKernelGGL(k_cell_partials, dim3((unsigned)((NC * 8 + 255) / 256)), dim3(256), 0, m->s, m->nx,
					   m->ny, m->nz, (const double *)m->gx, (const double *)m->gy, (const double *)m->gz,
					   (const int *)m->start, (const int *)m->cnt, (const MkRec *)m->prec, prm->nukoef,
					   prm->viscmod, m->part);
		hipLaunchKernelGGL(k_node_sums, gg, dim3(256), 0, m->s, m->nx, m->ny, m->nz, (const double *)m->part,
				   prm->nubeg, prm->nuend, prm->viscmod, m->fld);
	} else {
		hipLaunchKernelGGL(k_marker_gather, gg, dim3(256), 0, m->s, m->nx, m->ny, m->nz, (const double *)m->gx,
				   (const double *)m->gy, (const double *)m->gz, (const int *)m->start, (const int *)m->cnt,
				   (const int *)m->id, (const MkRec *)m->prec, prm->nukoef, prm->nubeg, prm->nuend,
				   prm->viscmod, m->fld);
	}
	if (m->has_tbc) {
		MCHK(hipMemcpyAsync(m->tk1, m->fld + 6 * N, N * sizeof(float), hipMemcpyDeviceToDevice, m->s));
		hipLaunchKernelGGL(k_tkbc, g1(N), dim3(256), 0, m->s, N, (const int *)m->tbc_p, (const float *)m->tbc_c,
				   (const float *)m->tbc_k, (const float *)m->tk1, m->fld + 6 * N);
	}
	MCHK(hipGetLastError());
	mark_end(m);
	int bad = 0;
	MCHK(hipMemcpyAsync(&bad, m->bad, sizeof(int), hipMemcpyDeviceToHost, m->s));
	MCHK(hipStreamSynchronize(m->s));
	if (bad) return I3G_EINVAL;
	mark_read(m);
	for (int q = 0; q < 11; q++)
		MCHK(hipMemcpyAsync(host[q], m->fld + q * N, N * sizeof(float), hipMemcpyDefault, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_set_tbc(i3g_markers *m, const int *bondm_t, const float *bondv, const int *bondn,
				   long nrec)
{
	if (!m) return I3G_EINVAL;
	if (!bondm_t) {
		m->has_tbc = 0;
		return I3G_OK;
	}
	if (!bondv || !bondn || nrec < 1) return I3G_EINVAL;
	long N = m->N;
	std::vector<int> p(N, -2);
	std::vector<float> c(N, 0.f), k(N, 0.f);
	for (long i = 0; i < N; i++) {
		long r = bondm_t[i];
		if (r <= 0) continue;
		if (r >= nrec) return I3G_EBC;
		long q = (long)bondn[r] - 1;
		if (q < -1 || q >= N) return I3G_EBC;
		p[i] = (int)q;
		c[i] = bondv[2 * r];
		k[i] = bondv[2 * r + 1];
	}
	int err = 0;
	if (!m->tbc_p) {
		m->tbc_p = mk_alloc<int>(N, &err); m->tbc_c = mk_alloc<float>(N, &err);
		m->tbc_k = mk_alloc<float>(N, &err); m->tk1 = mk_alloc<float>(N, &err);
		if (err) return err;
	}
	MCHK(hipMemcpyAsync(m->tbc_p, p.data(), N * sizeof(int), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->tbc_c, c.data(), N * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipMemcpyAsync(m->tbc_k, k.data(), N * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipStreamSynchronize(m->s));
	m->has_tbc = 1;
	return I3G_OK;
}

extern "C" int i3g_markers_update_tk(i3g_markers *m, const float *tk, const float *tk2, const float *kt,
				     const float *ro, const float *cp, double heatdif, double timestep, double zdeep,
				     double tdeep)
{
	if (!m || !tk || !tk2 || !kt || !ro || !cp || !(heatdif >= 0)) return I3G_EINVAL;
	if (heatdif != 0 && !m->sorted) return I3G_EINVAL;
	long N = m->N, n = m->n;
	int err = 0;
	if (!m->tnod) {
		m->tnod = mk_alloc<float>(5 * N, &err);
		m->sn = mk_alloc<double>(N, &err);
		m->trec = mk_alloc<TkRec>(m->cap, &err);
		m->tin = mk_alloc<uint8_t>(m->cap, &err);
		if (err) return err;
	}
	if (heatdif != 0 && (err = ensure_part(m))) return err;
	const float *src[5] = {tk, tk2, kt, ro, cp};
	for (int q = 0; q < 5; q++)
		MCHK(hipMemcpyAsync(m->tnod + q * N, src[q], N * sizeof(float), hipMemcpyDefault, m->s));
	const float *f = m->tnod;
	mark_begin(m);
	if (n)
		hipLaunchKernelGGL(k_tk_marker, g1(n), dim3(256), lines_bytes(m), m->s, n, (const float *)m->x, (const float *)m->y,
				   (const float *)m->z, (const signed char *)m->type, m->tk, (const double *)m->gx,
				   (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz, f, f + N, f + 2 * N,
				   f + 3 * N, f + 4 * N, heatdif, timestep, zdeep, tdeep, m->trec, m->tin);
	if (n && heatdif != 0) {
		dim3 gg((unsigned)(((long)m->nx * m->ny + 255) / 256), (unsigned)m->nz);
		if (m->order == 1) {
			long NC = (long)(m->nx - 1) * (m->ny - 1) * (m->nz - 1);
			hipLaunchKernelGGL(k_tk_cell_partials, dim3((unsigned)((NC * 8 + 255) / 256)), dim3(256), 0, m->s,
					   m->nx, m->ny, m->nz, (const double *)m->gx, (const double *)m->gy,
					   (const double *)m->gz, (const int *)m->start, (const int *)m->cnt,
					   (const TkRec *)m->trec, m->part);
			hipLaunchKernelGGL(k_tk_node_sums, gg, dim3(256), 0, m->s, m->nx, m->ny, m->nz,
					   (const double *)m->part, m->sn);
		} else {
			hipLaunchKernelGGL(k_tk_gather, gg, dim3(256), 0, m->s, m->nx, m->ny, m->nz, (const double *)m->gx,
					   (const double *)m->gy, (const double *)m->gz, (const int *)m->start,
					   (const int *)m->cnt, (const int *)m->id, (const TkRec *)m->trec, m->sn);
		}
		hipLaunchKernelGGL(k_tk_anti, g1(n), dim3(256), lines_bytes(m), m->s, n, (const float *)m->x, (const float *)m->y,
				   (const float *)m->z, (const uint8_t *)m->tin, m->tk, (const double *)m->gx,
				   (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz, (const double *)m->sn);
	}
	MCHK(hipGetLastError());
	mark_end(m);
	MCHK(hipStreamSynchronize(m->s));
	mark_read(m);
	return I3G_OK;
}

extern "C" int i3g_markers_advect_ex(i3g_markers *m, double dt, const float *vx, const float *vy, const float *vz,
				     const i3g_move_params *mp, const double *pr, const float *markd)
{
	if (!m || !vx || !vy || !vz || !mp || mp->outgrid < 0 || mp->outgrid > 2) return I3G_EINVAL;
	long N = m->N, n = m->n;
	int err = 0;
	if (!m->vel) {
		m->vel = mk_alloc<float>(3 * N, &err);
		if (err) return err;
	}
	MCHK(hipMemcpyAsync(m->vel, vx, N * sizeof(float), hipMemcpyDefault, m->s));
	MCHK(hipMemcpyAsync(m->vel + N, vy, N * sizeof(float), hipMemcpyDefault, m->s));
	MCHK(hipMemcpyAsync(m->vel + 2 * N, vz, N * sizeof(float), hipMemcpyDefault, m->s));
	MoveArgs A{*mp, nullptr, nullptr, (const int *)m->id};
	if (pr) {
		if (!m->pr) {
			m->pr = mk_alloc<double>(N, &err);
			if (err) return err;
		}
		MCHK(hipMemcpyAsync(m->pr, pr, N * sizeof(double), hipMemcpyDefault, m->s));
		A.pr = m->pr;
	}
	if (markd && n) {
		/* densities by marker id: as many as the largest id + 1 */
		std::vector<int> ids(n);
		MCHK(hipMemcpyAsync(ids.data(), m->id, n * sizeof(int), hipMemcpyDeviceToHost, m->s));
		MCHK(hipStreamSynchronize(m->s));
		long nid = 0;
		for (long i = 0; i < n; i++) nid = ids[i] + 1L > nid ? ids[i] + 1L : nid;
		if (nid > m->md_cap) {
			if (m->md) hipFree(m->md);
			m->md = mk_alloc<float>(nid, &err);
			if (err) { m->md = nullptr; m->md_cap = 0; return err; }
			m->md_cap = nid;
		}
		MCHK(hipMemcpyAsync(m->md, markd, nid * sizeof(float), hipMemcpyDefault, m->s));
		A.md = m->md;
	}
	mark_begin(m);
	if (n)
		hipLaunchKernelGGL(k_rk4, g1(n), dim3(256), lines_bytes(m), m->s, n, m->x, m->y, m->z, m->tk, m->eta,
				   m->type, (const float *)m->vel, (const float *)(m->vel + N),
				   (const float *)(m->vel + 2 * N), (const double *)m->gx, (const double *)m->gy,
				   (const double *)m->gz, m->nx, m->ny, m->nz, dt, DepthRules{m->zdeep, m->vdeep, m->tdeep}, A);
	MCHK(hipGetLastError());
	mark_end(m);
	MCHK(hipStreamSynchronize(m->s));
	mark_read(m);
	m->sorted = 0;
	return I3G_OK;
}

extern "C" int i3g_markers_advect_rk4(i3g_markers *m, double dt, const float *vx, const float *vy, const float *vz)
{
	i3g_move_params mp;
	memset(&mp, 0, sizeof(mp));
	return i3g_markers_advect_ex(m, dt, vx, vy, vz, &mp, nullptr, nullptr);
}

extern "C" int i3g_markers_set_grain(i3g_markers *m, const float *rr, const float *rn)
{
	if (!m || (!rr && !rn)) return I3G_EINVAL;
	if (m->n && rr) MCHK(hipMemcpyAsync(m->rr, rr, m->n * sizeof(float), hipMemcpyDefault, m->s));
	if (m->n && rn) MCHK(hipMemcpyAsync(m->rn, rn, m->n * sizeof(float), hipMemcpyDefault, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_set_melt(i3g_markers *m, long n, const float *w, const float *ex)
{
	if (!m || n < 0 || (n > 0 && (!w || !ex))) return I3G_EINVAL;
	if (n > m->mw_cap) {
		if (m->mw) hipFree(m->mw);
		m->mw = nullptr;
		m->mw_cap = m->mw_n = 0;
		int err = 0;
		m->mw = mk_alloc<float>(2 * n, &err);
		if (err) return err;
		m->mw_cap = n;
	}
	m->mw_n = n;
	if (n) {
		MCHK(hipMemcpyAsync(m->mw, w, n * sizeof(float), hipMemcpyDefault, m->s));
		MCHK(hipMemcpyAsync(m->mw + n, ex, n * sizeof(float), hipMemcpyDefault, m->s));
		MCHK(hipStreamSynchronize(m->s));
	}
	return I3G_OK;
}

extern "C" int i3g_markers_get_grain(i3g_markers *m, float *rr, float *rn)
{
	if (!m || (!rr && !rn)) return I3G_EINVAL;
	if (m->n && rr) MCHK(hipMemcpyAsync(rr, m->rr, m->n * sizeof(float), hipMemcpyDefault, m->s));
	if (m->n && rn) MCHK(hipMemcpyAsync(rn, m->rn, m->n * sizeof(float), hipMemcpyDefault, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_set_strain(i3g_markers *m, const float *e)
{
	if (!m || !e) return I3G_EINVAL;
	if (m->n) MCHK(hipMemcpyAsync(m->e, e, m->n * sizeof(float), hipMemcpyHostToDevice, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_get_strain(i3g_markers *m, float *e)
{
	if (!m || !e) return I3G_EINVAL;
	if (m->n) MCHK(hipMemcpyAsync(e, m->e, m->n * sizeof(float), hipMemcpyDeviceToHost, m->s));
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

/* the twelve strain-rate / stress planes and P into device buffers */
static int upload_st12(i3g_markers *m, const float *const *st12, const double *pr)
{
	long N = m->N;
	int err = 0;
	if (!m->st12) {
		m->st12 = mk_alloc<float>(12 * N, &err);
		if (err) return err;
	}
	if (!m->pr) {
		m->pr = mk_alloc<double>(N, &err);
		if (err) return err;
	}
	for (int q = 0; q < 12; q++) {
		if (!st12[q]) return I3G_EINVAL;
		MCHK(hipMemcpyAsync(m->st12 + q * N, st12[q], N * sizeof(float), hipMemcpyDefault, m->s));
	}
	m->has_st = 1;
	if (pr) MCHK(hipMemcpyAsync(m->pr, pr, N * sizeof(double), hipMemcpyDefault, m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_set_stress(i3g_markers *m, const float *const *st12)
{
	if (!m) return I3G_EINVAL;
	if (!st12) {
		m->has_st = 0;
		return I3G_OK;
	}
	int rc = upload_st12(m, st12, nullptr);
	if (rc) return rc;
	MCHK(hipStreamSynchronize(m->s));
	return I3G_OK;
}

extern "C" int i3g_markers_epscalc(i3g_markers *m, const float *const *st12, const double *pr, double *out)
{
	if (!m || !st12 || !pr || !out) return I3G_EINVAL;
	long n = m->n;
	int rc = upload_st12(m, st12, pr);
	if (rc) return rc;
	int err = 0;
	if (n) {
		double *d = mk_alloc<double>(14 * n, &err);
		if (err) return err;
		mark_begin(m);
		hipLaunchKernelGGL(k_epscalc, g1(n), dim3(256), lines_bytes(m), m->s, n, (const float *)m->x,
				   (const float *)m->y, (const float *)m->z, (const float *)m->st12, m->N, (const double *)m->pr,
				   (const double *)m->gx, (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz, d);
		mark_end(m);
		hipError_t e1 = hipGetLastError();
		hipError_t e2 = e1 == hipSuccess ? hipMemcpyAsync(out, d, 14 * n * sizeof(double), hipMemcpyDeviceToHost, m->s)
						 : e1;
		hipError_t e3 = hipStreamSynchronize(m->s);
		hipFree(d);
		if (e2 != hipSuccess || e3 != hipSuccess) return I3G_EDEVICE;
		mark_read(m);
	}
	return I3G_OK;
}

extern "C" int i3g_markers_update_props(i3g_markers *m, const i3g_rock *rocks, int nrock, const i3g_mk_params *prm,
					double healrate, const float *const *st12, const double *pr)
{
	if (!m || !rocks || !prm || !st12 || !pr || nrock < 1 || nrock > MAXROCK) return I3G_EINVAL;
	if (prm->outgrid < 0 || prm->outgrid > 2) return I3G_EINVAL;
	long n = m->n;
	int rc = upload_st12(m, st12, pr);
	if (rc) return rc;
	MCHK(hipMemcpyAsync(m->rocks, rocks, nrock * sizeof(i3g_rock), hipMemcpyDefault, m->s));
	MCHK(hipMemsetAsync(m->bad, 0, sizeof(int), m->s));
	mark_begin(m);
	if (n)
		hipLaunchKernelGGL(k_mk_props, g1(n), dim3(256), lines_bytes(m), m->s, n, (const float *)m->x,
				   (const float *)m->y, (const float *)m->z, (const float *)m->tk, m->eta, m->e, m->rr, m->rn,
				   (const signed char *)m->type, (const float *)m->st12, m->N, (const double *)m->pr,
				   (const double *)m->gx, (const double *)m->gy, (const double *)m->gz, m->nx, m->ny, m->nz,
				   (const i3g_rock *)m->rocks, nrock, *prm, healrate, m->bad);
	MCHK(hipGetLastError());
	mark_end(m);
	int bad = 0;
	MCHK(hipMemcpyAsync(&bad, m->bad, sizeof(int), hipMemcpyDeviceToHost, m->s));
	MCHK(hipStreamSynchronize(m->s));
	mark_read(m);
	return bad ? I3G_EINVAL : I3G_OK;
}

extern "C" int i3g_markers_set_depth_rules(i3g_markers *m, double zdeep, double vdeep, double tdeep)
{
	if (!m) return I3G_EINVAL;
	m->zdeep = zdeep;
	m->vdeep = vdeep;
	m->tdeep = tdeep;
	return I3G_OK;
}

/* ------------------------------------------------------------------ */
/* Rock bodies of the initial condition (in3mg.c:548-739)             */
/* ------------------------------------------------------------------ */

/* gausmat (gaus3mg.c:23-62): elimination without pivoting where each row
 * below the pivot is divided by its own leading coefficient before the
 * pivot row divided by the pivot is subtracted, then back substitution;
 * rm x (rm+1) augmented matrix, rows of rm+1 */
static void gauss_rows(int rm, double *a, double *sol)
{
	int rm1 = rm + 1;
	for (int m1 = 0; m1 < rm - 1; m1++) {
		double v1 = a[m1 * rm1 + m1];
		if (v1 == 0) continue;
		for (int m2 = m1 + 1; m2 < rm; m2++) {
			double v2 = a[m2 * rm1 + m1];
			if (v2 == 0) continue;
			for (int m3 = m1 + 1; m3 <= rm; m3++) a[m2 * rm1 + m3] = a[m2 * rm1 + m3] / v2 - a[m1 * rm1 + m3] / v1;
		}
	}
	for (int m1 = rm - 1; m1 >= 0; m1--) {
		sol[m1] = 0;
		double v1 = a[m1 * rm1 + m1];
		if (v1 != 0) {
			double v2 = a[m1 * rm1 + rm] / v1;
			sol[m1] = v2;
			for (int m2 = 0; m2 <= m1 - 1; m2++) {
				double w = a[m2 * rm1 + m1];
				if (w != 0) a[m2 * rm1 + rm] -= w * v2;
			}
		}
	}
}

/* one bilinear face through the four corners c[]: value v = a0 + a1 p + a2
 * q + a3 p q, with (p, q) the other two coordinates in the order the
 * reference writes the matrix rows */
static void face_coefs(const double *p, const double *q, const double *v, const int *c, double *out)
{
	double mat[20];
	for (int r = 0; r < 4; r++) {
		int i = c[r];
		mat[r * 5 + 0] = 1.0;
		mat[r * 5 + 1] = p[i];
		mat[r * 5 + 2] = q[i];
		mat[r * 5 + 3] = p[i] * q[i];
		mat[r * 5 + 4] = v[i];
	}
	gauss_rows(4, mat, out);
}

struct PaintBox {
	int shape, type;
	float ex, e;
	double k[24]; /* hexahedron: left, right (x), top, bottom (y), front, back (z) faces */
	double c[8];  /* shells: centre x y z, inner, outer, scales x y z */
};

__global__ void k_paint(long n, const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
			float *tk, float *e, signed char *type, float *ex, float *od, float *ow, const PaintBox *B, int nbox,
			double xs, double ys, double zs)
{
#pragma clang fp contract(off)
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	float fx = x[t], fy = y[t], fz = z[t];
	double px = fx / xs, py = fy / ys, pz = fz / zs;
	int hit = -1, hex = -1;
	for (int b = 0; b < nbox; b++) {
		const PaintBox &P = B[b];
		bool in = false;
		if (P.shape == 0) {
			const double *k = P.k;
			in = px >= k[0] + k[1] * py + k[2] * pz + k[3] * py * pz && px <= k[4] + k[5] * py + k[6] * pz + k[7] * py * pz &&
			     py >= k[8] + k[9] * px + k[10] * pz + k[11] * px * pz &&
			     py <= k[12] + k[13] * px + k[14] * pz + k[15] * px * pz &&
			     pz >= k[16] + k[17] * py + k[18] * px + k[19] * py * px &&
			     pz <= k[20] + k[21] * py + k[22] * px + k[23] * py * px;
		} else {
			double dx = (px - P.c[0]) / P.c[5];
			double dy = (py - P.c[1]) / P.c[6];
			double dz = (pz - P.c[2]) / P.c[7];
			if (P.shape == 1) {
				double ival = pow(dx * dx + dy * dy + dz * dz, 0.5);
				in = ival >= P.c[3] && ival <= P.c[4];
			} else {
				double ival = pow(dx * dx + dz * dz, 0.5);
				in = ival >= P.c[3] && ival <= P.c[4] && fabs(dy) <= 1.0;
			}
		}
		if (in) {
			hit = b;
			if (P.shape == 0) hex = b;
		}
	}
	if (hit < 0) return;
	type[t] = (signed char)B[hit].type;
	e[t] = B[hit].e;
	if (ex) ex[t] = B[hit].ex;
	if (hex >= 0) {
		/* the last hexahedron sets markk/markd/markw; a later shell does not
		 * touch them */
		bool imm = B[hex].type >= 100;
		tk[t] = imm ? fx : 0.0f;
		if (od) od[t] = imm ? fy : 0.0f;
		if (ow) ow[t] = imm ? fz : 0.0f;
	}
}

extern "C" int i3g_markers_paint(i3g_markers *m, double xsize, double ysize, double zsize, const i3g_rock_box *boxes,
				 int nbox, float *ex, float *od, float *ow)
{
	if (!m || !boxes || nbox < 1 || !(xsize > 0) || !(ysize > 0) || !(zsize > 0)) return I3G_EINVAL;
	std::vector<PaintBox> pb(nbox);
	for (int b = 0; b < nbox; b++) {
		const i3g_rock_box &r = boxes[b];
		if (r.shape < 0 || r.shape > 2 || r.type < -128 || r.type > 127) return I3G_EINVAL;
		PaintBox &P = pb[b];
		memset(&P, 0, sizeof(P));
		P.shape = r.shape;
		P.type = r.type;
		P.ex = (float)r.ex;
		P.e = (float)r.e;
		if (r.shape == 0) {
			/* the six faces in the order and with the corners of
			 * in3mg.c:592-694 */
			const int left[4] = {0, 1, 4, 5}, right[4] = {2, 3, 6, 7}, top[4] = {0, 2, 4, 6}, bot[4] = {1, 3, 5, 7};
			const int front[4] = {0, 1, 2, 3}, back[4] = {4, 5, 6, 7};
			face_coefs(r.y, r.z, r.x, left, P.k + 0);
			face_coefs(r.y, r.z, r.x, right, P.k + 4);
			face_coefs(r.x, r.z, r.y, top, P.k + 8);
			face_coefs(r.x, r.z, r.y, bot, P.k + 12);
			face_coefs(r.y, r.x, r.z, front, P.k + 16);
			face_coefs(r.y, r.x, r.z, back, P.k + 20);
		} else {
			P.c[0] = r.x[0]; P.c[1] = r.y[0]; P.c[2] = r.z[0];
			P.c[3] = r.x[2]; P.c[4] = r.x[5];
			P.c[5] = r.x[7]; P.c[6] = r.y[7]; P.c[7] = r.z[7];
		}
	}
	long n = m->n;
	int err = 0;
	PaintBox *dB = mk_alloc<PaintBox>(nbox, &err);
	float *dex = ex ? mk_alloc<float>(n, &err) : nullptr;
	float *dod = od ? mk_alloc<float>(n, &err) : nullptr;
	float *dow = ow ? mk_alloc<float>(n, &err) : nullptr;
	if (err) {
		hipFree(dB); hipFree(dex); hipFree(dod); hipFree(dow);
		return err;
	}
	int rc = I3G_OK;
	hipError_t e = hipMemcpyAsync(dB, pb.data(), nbox * sizeof(PaintBox), hipMemcpyHostToDevice, m->s);
	if (e == hipSuccess && ex && n) e = hipMemcpyAsync(dex, ex, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess && od && n) e = hipMemcpyAsync(dod, od, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess && ow && n) e = hipMemcpyAsync(dow, ow, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess && n) {
		mark_begin(m);
		hipLaunchKernelGGL(k_paint, g1(n), dim3(256), 0, m->s, n, (const float *)m->x, (const float *)m->y,
				   (const float *)m->z, m->tk, m->e, m->type, dex, dod, dow, (const PaintBox *)dB, nbox, xsize,
				   ysize, zsize);
		e = hipGetLastError();
		mark_end(m);
	}
	if (e == hipSuccess && ex && n) e = hipMemcpyAsync(ex, dex, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess && od && n) e = hipMemcpyAsync(od, dod, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess && ow && n) e = hipMemcpyAsync(ow, dow, n * sizeof(float), hipMemcpyDefault, m->s);
	if (e == hipSuccess) e = hipStreamSynchronize(m->s);
	if (e != hipSuccess) rc = I3G_EDEVICE;
	else mark_read(m);
	hipFree(dB); hipFree(dex); hipFree(dod); hipFree(dow);
	return rc;
}
