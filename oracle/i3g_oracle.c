/*
 * i3g oracle — TEST INFRASTRUCTURE ONLY (see i3g_oracle.h).
 * Plain-C FP64 restatement of the reference Stokes operator, transfers and
 * norms.  Each function cites the reference lines it follows.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "i3g_oracle.h"

/* Node sums of the marker -> node scatters: 0 in float, marker by marker,
 * as the reference's float node arrays (its serial order); 1 in double
 * (the exact sum rounded once, the value a reordered FP64 reduction
 * approaches). */
static int or_sum64 = 0;
void or_set_sum64(int on) { or_sum64 = on; }

/* one scatter addition into node accumulator k (float or double) */
static inline void or_acc_add(float *acc, double *accd, long k, double v)
{
	if (or_sum64) accd[k] += v;
	else acc[k] = (float)(acc[k] + v);
}

static inline double or_acc_get(const float *acc, const double *accd, long k)
{
	return or_sum64 ? accd[k] : (double)acc[k];
}

#define NODE(L, m1, m2, m3) ((m3) * (L)->nx * (L)->ny + (m1) * (L)->ny + (m2))

static double clampd(double v, double lo, double hi)
{
	if (v > hi) v = hi;
	if (v < lo) v = lo;
	return v;
}

static double avg8(const float *nu, const long *v, int viscmod)
{
	int k;
	double s = 0;
	if (viscmod == 1) {
		for (k = 0; k < 8; k++) s += log(nu[v[k]]);
		return exp(s / 8.0);
	}
	if (viscmod == 2) {
		for (k = 0; k < 8; k++) s += 1.0 / nu[v[k]];
		return 1.0 / (s / 8.0);
	}
	{
		/* the reference adds the eight float values in float */
		float fs = nu[v[0]];
		for (k = 1; k < 8; k++) fs = fs + nu[v[k]];
		s = fs;
	}
	return s / 8.0;
}

static double avg2(float a, float b, int viscmod)
{
	if (viscmod == 1) return exp((log(a) + log(b)) / 2.0);
	if (viscmod == 2) return 1.0 / ((1.0 / a + 1.0 / b) / 2.0);
	return (double)(float)(a + b) / 2.0;
}

/* move3mg.c:36-48 (cells), 402-414 (xy), 517-529 (xz), 632-644 (yz) */
void or_eta(const orl_t *L, const float *nu, const float *nxx, const float *nxy,
	    const float *nxz, const float *nyz, int viscmod, double nubeg,
	    double nuend, double *etaC, double *etaXY, double *etaXZ,
	    double *etaYZ)
{
	long nx = L->nx, ny = L->ny, nz = L->nz, m1, m2, m3;
	int lev0 = L->level == 0;
	for (m3 = 0; m3 < nz; m3++)
	for (m1 = 0; m1 < nx; m1++)
	for (m2 = 0; m2 < ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		double e;
		etaC[n] = etaXY[n] = etaXZ[n] = etaYZ[n] = 0;
		if (m1 >= 1 && m2 >= 1 && m3 >= 1) {
			if (lev0 && m1 > 3 && m2 > 3 && m3 > 3 && m1 < nx - 3 &&
			    m2 < ny - 3 && m3 < nz - 3) {
				e = nxx[n];
			} else {
				long v[8];
				v[0] = NODE(L, m1 - 1, m2 - 1, m3 - 1); v[1] = v[0] + 1;
				v[2] = v[0] + ny; v[3] = v[2] + 1;
				v[4] = v[0] + nx * ny; v[5] = v[4] + 1;
				v[6] = v[4] + ny; v[7] = v[6] + 1;
				e = avg8(nu, v, viscmod);
			}
			etaC[n] = clampd(e, nubeg, nuend);
		}
		if (m3 < nz - 1) {
			if (lev0 && m1 > 3 && m2 > 3 && m3 > 2 && m1 < nx - 4 &&
			    m2 < ny - 4 && m3 < nz - 4)
				e = nxy[n];
			else
				e = avg2(nu[n], nu[n + nx * ny], viscmod);
			etaXY[n] = clampd(e, nubeg, nuend);
		}
		if (m2 < ny - 1) {
			if (lev0 && m1 > 3 && m2 > 2 && m3 > 3 && m1 < nx - 4 &&
			    m2 < ny - 4 && m3 < nz - 4)
				e = nxz[n];
			else
				e = avg2(nu[n], nu[n + 1], viscmod);
			etaXZ[n] = clampd(e, nubeg, nuend);
		}
		if (m1 < nx - 1) {
			if (lev0 && m1 > 2 && m2 > 3 && m3 > 3 && m1 < nx - 4 &&
			    m2 < ny - 4 && m3 < nz - 4)
				e = nyz[n];
			else
				e = avg2(nu[n], nu[n + ny], viscmod);
			etaYZ[n] = clampd(e, nubeg, nuend);
		}
	}
}

static int exists(const orl_t *L, long m1, long m2, long m3, int c)
{
	switch (c) {
	case 0: return m1 && m2 && m3;
	case 1: return m2 < L->ny - 1 && m3 < L->nz - 1;
	case 2: return m1 < L->nx - 1 && m3 < L->nz - 1;
	default: return m1 < L->nx - 1 && m2 < L->ny - 1;
	}
}

/* move3mg.c:751, :943, :1142 — RHS = -g_i * mean of 4 node densities */
void or_rhs(const orl_t *L, const float *ro, const double *g, const int *bondm,
	    const float *bondv, double *b)
{
	long nx = L->nx, ny = L->ny, nz = L->nz, m1, m2, m3, sx = ny, sz = nx * ny;
	int c;
	for (m3 = 0; m3 < nz; m3++)
	for (m1 = 0; m1 < nx; m1++)
	for (m2 = 0; m2 < ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		for (c = 0; c < 4; c++) {
			long u = n * 4 + c;
			int rec = bondm[(L->mgp + n) * 4 + c];
			b[u] = 0;
			if (!exists(L, m1, m2, m3, c)) continue;
			if (rec) { b[u] = bondv[2 * rec]; continue; }
			if (c == 1)
				b[u] = -g[0] * (ro[n] + ro[n + 1] + ro[n + sz] + ro[n + 1 + sz]) / 4.0;
			else if (c == 2)
				b[u] = -g[1] * (ro[n] + ro[n + sx] + ro[n + sz] + ro[n + sx + sz]) / 4.0;
			else if (c == 3)
				b[u] = -g[2] * (ro[n] + ro[n + 1] + ro[n + sx] + ro[n + sx + 1]) / 4.0;
		}
	}
}

/* Stresses in terms of x (interleaved) — value forms of sxxcalc..syzcalc
 * (move3mg.c:3-710); eta from or_eta. */
typedef struct {
	const orl_t *L;
	const double *eC, *eXY, *eXZ, *eYZ, *x;
} ctx_t;

#define XV(C, m1, m2, m3, c) ((C)->x[NODE((C)->L, m1, m2, m3) * 4 + (c)])

static double s_xx(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eC[NODE(L, a, b, c)] *
	       (XV(C, a, b - 1, c - 1, 1) - XV(C, a - 1, b - 1, c - 1, 1)) /
	       (L->gx[a] - L->gx[a - 1]);
}
static double s_yy(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eC[NODE(L, a, b, c)] *
	       (XV(C, a - 1, b, c - 1, 2) - XV(C, a - 1, b - 1, c - 1, 2)) /
	       (L->gy[b] - L->gy[b - 1]);
}
static double s_zz(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eC[NODE(L, a, b, c)] *
	       (XV(C, a - 1, b - 1, c, 3) - XV(C, a - 1, b - 1, c - 1, 3)) /
	       (L->gz[c] - L->gz[c - 1]);
}
static double s_xy(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eXY[NODE(L, a, b, c)] *
	       ((XV(C, a, b, c, 1) - XV(C, a, b - 1, c, 1)) / (L->gy[b + 1] - L->gy[b - 1]) +
		(XV(C, a, b, c, 2) - XV(C, a - 1, b, c, 2)) / (L->gx[a + 1] - L->gx[a - 1]));
}
static double s_xz(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eXZ[NODE(L, a, b, c)] *
	       ((XV(C, a, b, c, 1) - XV(C, a, b, c - 1, 1)) / (L->gz[c + 1] - L->gz[c - 1]) +
		(XV(C, a, b, c, 3) - XV(C, a - 1, b, c, 3)) / (L->gx[a + 1] - L->gx[a - 1]));
}
static double s_yz(const ctx_t *C, long a, long b, long c)
{
	const orl_t *L = C->L;
	return 2.0 * C->eYZ[NODE(L, a, b, c)] *
	       ((XV(C, a, b, c, 2) - XV(C, a, b, c - 1, 2)) / (L->gz[c + 1] - L->gz[c - 1]) +
		(XV(C, a, b, c, 3) - XV(C, a, b - 1, c, 3)) / (L->gy[b + 1] - L->gy[b - 1]));
}

/* A x for a stencil row (x/y/zstokserr move3mg.c:718-1301, continerr
 * :1310-1401) and the row's diagonal coefficient. */
static double row_ax(const ctx_t *C, long i, long j, long k, int c, double *diag)
{
	const orl_t *L = C->L;
	const double *gx = L->gx, *gy = L->gy, *gz = L->gz;
	double ax = 0, xkf, ykf, zkf, d = 0;
	long nyx = L->ny, nxy = L->nx * L->ny;
	if (c == 1) {
		xkf = 2.0 / (gx[i + 1] - gx[i - 1]);
		ykf = 1.0 / (gy[j + 1] - gy[j]);
		zkf = 1.0 / (gz[k + 1] - gz[k]);
		ax = xkf * (s_xx(C, i + 1, j + 1, k + 1) - s_xx(C, i, j + 1, k + 1)) +
		     ykf * (s_xy(C, i, j + 1, k) - s_xy(C, i, j, k)) +
		     zkf * (s_xz(C, i, j, k + 1) - s_xz(C, i, j, k)) -
		     xkf * (XV(C, i + 1, j + 1, k + 1, 0) - XV(C, i, j + 1, k + 1, 0));
		d = -xkf * (2.0 * C->eC[NODE(L, i + 1, j + 1, k + 1)] / (gx[i + 1] - gx[i]) +
			    2.0 * C->eC[NODE(L, i, j + 1, k + 1)] / (gx[i] - gx[i - 1])) -
		    ykf * (2.0 * C->eXY[NODE(L, i, j + 1, k)] / (gy[j + 2] - gy[j]) +
			   2.0 * C->eXY[NODE(L, i, j, k)] / (gy[j + 1] - gy[j - 1])) -
		    zkf * (2.0 * C->eXZ[NODE(L, i, j, k + 1)] / (gz[k + 2] - gz[k]) +
			   2.0 * C->eXZ[NODE(L, i, j, k)] / (gz[k + 1] - gz[k - 1]));
	} else if (c == 2) {
		xkf = 1.0 / (gx[i + 1] - gx[i]);
		ykf = 2.0 / (gy[j + 1] - gy[j - 1]);
		zkf = 1.0 / (gz[k + 1] - gz[k]);
		ax = ykf * (s_yy(C, i + 1, j + 1, k + 1) - s_yy(C, i + 1, j, k + 1)) +
		     xkf * (s_xy(C, i + 1, j, k) - s_xy(C, i, j, k)) +
		     zkf * (s_yz(C, i, j, k + 1) - s_yz(C, i, j, k)) -
		     ykf * (XV(C, i + 1, j + 1, k + 1, 0) - XV(C, i + 1, j, k + 1, 0));
		d = -ykf * (2.0 * C->eC[NODE(L, i + 1, j + 1, k + 1)] / (gy[j + 1] - gy[j]) +
			    2.0 * C->eC[NODE(L, i + 1, j, k + 1)] / (gy[j] - gy[j - 1])) -
		    xkf * (2.0 * C->eXY[NODE(L, i + 1, j, k)] / (gx[i + 2] - gx[i]) +
			   2.0 * C->eXY[NODE(L, i, j, k)] / (gx[i + 1] - gx[i - 1])) -
		    zkf * (2.0 * C->eYZ[NODE(L, i, j, k + 1)] / (gz[k + 2] - gz[k]) +
			   2.0 * C->eYZ[NODE(L, i, j, k)] / (gz[k + 1] - gz[k - 1]));
	} else if (c == 3) {
		xkf = 1.0 / (gx[i + 1] - gx[i]);
		ykf = 1.0 / (gy[j + 1] - gy[j]);
		zkf = 2.0 / (gz[k + 1] - gz[k - 1]);
		ax = zkf * (s_zz(C, i + 1, j + 1, k + 1) - s_zz(C, i + 1, j + 1, k)) +
		     xkf * (s_xz(C, i + 1, j, k) - s_xz(C, i, j, k)) +
		     ykf * (s_yz(C, i, j + 1, k) - s_yz(C, i, j, k)) -
		     zkf * (XV(C, i + 1, j + 1, k + 1, 0) - XV(C, i + 1, j + 1, k, 0));
		d = -zkf * (2.0 * C->eC[NODE(L, i + 1, j + 1, k + 1)] / (gz[k + 1] - gz[k]) +
			    2.0 * C->eC[NODE(L, i + 1, j + 1, k)] / (gz[k] - gz[k - 1])) -
		    xkf * (2.0 * C->eXZ[NODE(L, i + 1, j, k)] / (gx[i + 2] - gx[i]) +
			   2.0 * C->eXZ[NODE(L, i, j, k)] / (gx[i + 1] - gx[i - 1])) -
		    ykf * (2.0 * C->eYZ[NODE(L, i, j + 1, k)] / (gy[j + 2] - gy[j]) +
			   2.0 * C->eYZ[NODE(L, i, j, k)] / (gy[j + 1] - gy[j - 1]));
	} else {
		ax = (XV(C, i, j - 1, k - 1, 1) - XV(C, i - 1, j - 1, k - 1, 1)) / (gx[i] - gx[i - 1]) +
		     (XV(C, i - 1, j, k - 1, 2) - XV(C, i - 1, j - 1, k - 1, 2)) / (gy[j] - gy[j - 1]) +
		     (XV(C, i - 1, j - 1, k, 3) - XV(C, i - 1, j - 1, k - 1, 3)) / (gz[k] - gz[k - 1]);
		d = 0;
	}
	(void)nyx; (void)nxy;
	if (diag) *diag = d;
	return ax;
}

/* bonderr (move3mg.c:1532-1579): A x = x_i - k x_partner, b = const. */
static double bc_ax(const orl_t *L, const int *bondn, const float *bondv, int rec,
		    long u, const double *x)
{
	double ax = x[u];
	if (bondn[rec]) ax -= bondv[2 * rec + 1] * x[bondn[rec] - 1 - L->mgp * 4];
	return ax;
}

void or_residual(const orl_t *L, const double *etaC, const double *etaXY,
		 const double *etaXZ, const double *etaYZ, const int *bondm,
		 const float *bondv, const int *bondn, const double *b,
		 const double *x, double *r)
{
	ctx_t C = {L, etaC, etaXY, etaXZ, etaYZ, x};
	long m1, m2, m3;
	int c;
	for (m3 = 0; m3 < L->nz; m3++)
	for (m1 = 0; m1 < L->nx; m1++)
	for (m2 = 0; m2 < L->ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		for (c = 0; c < 4; c++) {
			long u = n * 4 + c;
			int rec = bondm[(L->mgp + n) * 4 + c];
			r[u] = 0;
			if (!exists(L, m1, m2, m3, c)) continue;
			if (rec) r[u] = b[u] - bc_ax(L, bondn, bondv, rec, u, x);
			else r[u] = b[u] - row_ax(&C, m1, m2, m3, c, NULL);
		}
	}
}

void or_wnorms(const orl_t *L, const double *etaC, const int *bondm,
	       const double *b, const double *r, double *out)
{
	long m1, m2, m3;
	int c;
	double sr = 0, sb = 0;
	for (m3 = 0; m3 < L->nz; m3++)
	for (m1 = 0; m1 < L->nx; m1++)
	for (m2 = 0; m2 < L->ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		for (c = 0; c < 4; c++) {
			long u = n * 4 + c;
			double w = 1.0;
			if (!exists(L, m1, m2, m3, c)) continue;
			if (bondm[(L->mgp + n) * 4 + c]) continue;
			if (c == 0) {
				double h = ((L->gx[m1] - L->gx[m1 - 1]) + (L->gy[m2] - L->gy[m2 - 1]) +
					    (L->gz[m3] - L->gz[m3 - 1])) / 3.0;
				w = etaC[n] / h;
			}
			sr += (r[u] * w) * (r[u] * w);
			sb += (b[u] * w) * (b[u] * w);
		}
	}
	out[0] = sqrt(sr);
	out[1] = sqrt(sb);
}

/* move3mg.c:1691-1764: sigma/eps at cells (m1,m2,m3>0) and edges. */
void or_stress(const orl_t *L, const double *etaC, const double *etaXY,
	       const double *etaXZ, const double *etaYZ, const float *vx,
	       const float *vy, const float *vz, float *eps6, float *sig6)
{
	long nx = L->nx, ny = L->ny, nz = L->nz, N = nx * ny * nz, m1, m2, m3;
	const double *gx = L->gx, *gy = L->gy, *gz = L->gz;
	for (m3 = 0; m3 < nz; m3++)
	for (m1 = 0; m1 < nx; m1++)
	for (m2 = 0; m2 < ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		double e;
		if (m1 > 0 && m2 > 0 && m3 > 0) {
			long v0 = NODE(L, m1 - 1, m2 - 1, m3 - 1);
			e = (vx[v0 + ny] - vx[v0]) * (1.0 / (gx[m1] - gx[m1 - 1]));
			eps6[0 * N + n] = (float)e; sig6[0 * N + n] = (float)(2.0 * etaC[n] * e);
			e = (vy[v0 + 1] - vy[v0]) * (1.0 / (gy[m2] - gy[m2 - 1]));
			eps6[1 * N + n] = (float)e; sig6[1 * N + n] = (float)(2.0 * etaC[n] * e);
			e = (vz[v0 + nx * ny] - vz[v0]) * (1.0 / (gz[m3] - gz[m3 - 1]));
			eps6[2 * N + n] = (float)e; sig6[2 * N + n] = (float)(2.0 * etaC[n] * e);
		}
		if (m1 > 0 && m1 < nx - 1 && m2 > 0 && m2 < ny - 1 && m3 < nz - 1) {
			e = (vx[n] - vx[n - 1]) * (1.0 / (gy[m2 + 1] - gy[m2 - 1])) +
			    (vy[n] - vy[n - ny]) * (1.0 / (gx[m1 + 1] - gx[m1 - 1]));
			eps6[3 * N + n] = (float)e; sig6[3 * N + n] = (float)(2.0 * etaXY[n] * e);
		}
		if (m1 > 0 && m1 < nx - 1 && m2 < ny - 1 && m3 > 0 && m3 < nz - 1) {
			e = (vx[n] - vx[n - nx * ny]) * (1.0 / (gz[m3 + 1] - gz[m3 - 1])) +
			    (vz[n] - vz[n - ny]) * (1.0 / (gx[m1 + 1] - gx[m1 - 1]));
			eps6[4 * N + n] = (float)e; sig6[4 * N + n] = (float)(2.0 * etaXZ[n] * e);
		}
		if (m1 < nx - 1 && m2 > 0 && m2 < ny - 1 && m3 > 0 && m3 < nz - 1) {
			e = (vy[n] - vy[n - nx * ny]) * (1.0 / (gz[m3 + 1] - gz[m3 - 1])) +
			    (vz[n] - vz[n - 1]) * (1.0 / (gy[m2 + 1] - gy[m2 - 1]));
			eps6[5 * N + n] = (float)e; sig6[5 * N + n] = (float)(2.0 * etaYZ[n] * e);
		}
	}
}

/* ------------------------------------------------------------------ */
/* Transfers                                                          */
/* ------------------------------------------------------------------ */

/* Base coarse index of a fine index along one axis (the m10min/m20min/
 * m30min blocks of move3mg.c:2779-2789, 3201-3214): kind 0 = node line
 * (clamp n-2), 1 = mid-point (clamp n-3, shift down if the fine mid lies
 * below the coarse mid), 2 = cell centre (clamp n-2, shift if below). */
static long base_idx(long m, long nf, long nc, const double *gf, const double *gc,
		     int kind)
{
	long r = (m - 2) / 2 + 2;
	if (m < 3) r = m;
	if (m > nf - 4) r = nc - (nf - m);
	if (kind == 0) {
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

/* one-dimensional weight for coarse index cc (cc == base or base+1) */
static double w1d(long m, long base, long cc, const double *gf, const double *gc,
		  int kind)
{
	double d, dd;
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

/* move3mg.c:2642-2741 */
void or_ronurestrict(const orl_t *F, const orl_t *C, const float *nu_f,
		     float *nu_c, int viscmod)
{
	long Nc = C->nx * C->ny * C->nz, m1, m2, m3, n;
	double *sw = calloc((size_t)Nc, sizeof(double));
	for (n = 0; n < Nc; n++) nu_c[n] = 0;
	for (m3 = 0; m3 < F->nz; m3++)
	for (m1 = 0; m1 < F->nx; m1++)
	for (m2 = 0; m2 < F->ny; m2++) {
		long m4 = NODE(F, m1, m2, m3), a, b, c;
		long b1 = base_idx(m1, F->nx, C->nx, F->gx, C->gx, 0);
		long b2 = base_idx(m2, F->ny, C->ny, F->gy, C->gy, 0);
		long b3 = base_idx(m3, F->nz, C->nz, F->gz, C->gz, 0);
		double dx = C->gx[b1 + 1] - C->gx[b1];
		double dy = C->gy[b2 + 1] - C->gy[b2];
		double dz = C->gz[b3 + 1] - C->gz[b3];
		for (c = b3; c <= b3 + 1; c++) {
			double ddz = 1.0 - fabs((C->gz[c] - F->gz[m3]) / dz);
			for (a = b1; a <= b1 + 1; a++) {
				double ddx = 1.0 - fabs((C->gx[a] - F->gx[m1]) / dx);
				for (b = b2; b <= b2 + 1; b++) {
					double ddy = 1.0 - fabs((C->gy[b] - F->gy[m2]) / dy);
					long m40 = NODE(C, a, b, c);
					double swt = ddx * ddy * ddz, v;
					if (viscmod == 1) v = log(nu_f[m4]) * swt;
					else if (viscmod == 2) v = 1.0 / nu_f[m4] * swt;
					else v = nu_f[m4] * swt;
					nu_c[m40] = (float)(nu_c[m40] + v);
					sw[m40] += swt;
				}
			}
		}
	}
	for (n = 0; n < Nc; n++) {
		nu_c[n] = (float)(nu_c[n] / sw[n]);
		if (viscmod == 1) nu_c[n] = (float)exp(nu_c[n]);
		if (viscmod == 2) nu_c[n] = (float)(1.0 / nu_c[n]);
	}
	free(sw);
}

/* per-component axis kinds: [comp][axis] 0 node,1 mid,2 centre */
static const int KIND[4][3] = {{2, 2, 2}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

static void bases(const orl_t *F, const orl_t *C, int comp, long m1, long m2,
		  long m3, long *b)
{
	b[0] = base_idx(m1, F->nx, C->nx, F->gx, C->gx, KIND[comp][0]);
	b[1] = base_idx(m2, F->ny, C->ny, F->gy, C->gy, KIND[comp][1]);
	b[2] = base_idx(m3, F->nz, C->nz, F->gz, C->gz, KIND[comp][2]);
}

/* move3mg.c:3165-3520 */
void or_vprestrict(const orl_t *F, const orl_t *C, const int *bondm,
		   const double *r_f, const float *nd_f, const float *nd_c,
		   double *b_c)
{
	long Nc = C->nx * C->ny * C->nz, m1, m2, m3, u;
	double *sw = calloc((size_t)(4 * Nc), sizeof(double));
	int comp;
	for (u = 0; u < 4 * Nc; u++) b_c[u] = 0;
	for (m3 = 0; m3 < F->nz; m3++)
	for (m1 = 0; m1 < F->nx; m1++)
	for (m2 = 0; m2 < F->ny; m2++) {
		long m4 = NODE(F, m1, m2, m3);
		for (comp = 1; comp <= 4; comp++) {
			int cc = comp & 3; /* order vx, vy, vz, P */
			long b[3], a, bb, c;
			if (!exists(F, m1, m2, m3, cc)) continue;
			if (bondm[(F->mgp + m4) * 4 + cc]) continue;
			bases(F, C, cc, m1, m2, m3, b);
			for (c = b[2]; c <= b[2] + 1; c++) {
				double ddz = w1d(m3, b[2], c, F->gz, C->gz, KIND[cc][2]);
				for (a = b[0]; a <= b[0] + 1; a++) {
					double ddx = w1d(m1, b[0], a, F->gx, C->gx, KIND[cc][0]);
					for (bb = b[1]; bb <= b[1] + 1; bb++) {
						double ddy = w1d(m2, b[1], bb, F->gy, C->gy, KIND[cc][1]);
						long m40 = NODE(C, a, bb, c);
						double swt;
						if (bondm[(C->mgp + m40) * 4 + cc]) continue;
						swt = ddx * ddy * ddz;
						if (cc == 0)
							b_c[m40 * 4] += r_f[m4 * 4] * nd_f[m4] / nd_c[m40] * swt;
						else
							b_c[m40 * 4 + cc] += r_f[m4 * 4 + cc] * swt;
						sw[m40 * 4 + cc] += swt;
					}
				}
			}
		}
	}
	for (u = 0; u < 4 * Nc; u++)
		if (sw[u] != 0) b_c[u] /= sw[u];
	free(sw);
}

/* move3mg.c:2748-3063 */
void or_vpprolong(const orl_t *F, const orl_t *C, const double *x_c,
		  double *x_f, double v1koef, double p1koef)
{
	long m1, m2, m3;
	int comp;
	for (m3 = 0; m3 < F->nz; m3++)
	for (m1 = 0; m1 < F->nx; m1++)
	for (m2 = 0; m2 < F->ny; m2++) {
		long m4 = NODE(F, m1, m2, m3);
		for (comp = 1; comp <= 4; comp++) {
			int cc = comp & 3;
			long b[3], a, bb, c;
			double kf = cc ? v1koef : p1koef;
			if (!exists(F, m1, m2, m3, cc)) continue;
			bases(F, C, cc, m1, m2, m3, b);
			for (c = b[2]; c <= b[2] + 1; c++) {
				double ddz = w1d(m3, b[2], c, F->gz, C->gz, KIND[cc][2]);
				for (a = b[0]; a <= b[0] + 1; a++) {
					double ddx = w1d(m1, b[0], a, F->gx, C->gx, KIND[cc][0]);
					for (bb = b[1]; bb <= b[1] + 1; bb++) {
						double ddy = w1d(m2, b[1], bb, F->gy, C->gy, KIND[cc][1]);
						double swt = ddx * ddy * ddz;
						x_f[m4 * 4 + cc] += x_c[NODE(C, a, bb, c) * 4 + cc] * swt * kf;
					}
				}
			}
		}
	}
}

/* gausmat4 update rules (move3mg.c:3087-3131), lexicographic order. */
void or_gs_sweep(const orl_t *L, const double *etaC, const double *etaXY,
		 const double *etaXZ, const double *etaYZ, const int *bondm,
		 const float *bondv, const int *bondn, const double *b, double *x,
		 double p0koef, double v0koef)
{
	ctx_t C = {L, etaC, etaXY, etaXZ, etaYZ, x};
	long m1, m2, m3;
	int c;
	for (m3 = 0; m3 < L->nz; m3++)
	for (m1 = 0; m1 < L->nx; m1++)
	for (m2 = 0; m2 < L->ny; m2++) {
		long n = NODE(L, m1, m2, m3);
		for (c = 0; c < 4; c++) {
			long u = n * 4 + c;
			int rec = bondm[(L->mgp + n) * 4 + c];
			double ax, d, res;
			if (!exists(L, m1, m2, m3, c)) continue;
			if (rec) {
				res = bc_ax(L, bondn, bondv, rec, u, x) - b[u];
				x[u] -= res;
				continue;
			}
			ax = row_ax(&C, m1, m2, m3, c, &d);
			res = ax - b[u];
			if (c == 0) x[u] -= res * etaC[n] * p0koef;
			else x[u] -= res / d * v0koef;
		}
	}
}

/* One phase of the colour-ordered sweep with the gausmat4 update rules
 * (move3mg.c:3100-3128): phase 0..5 relaxes velocity component 1 + phase/2
 * on the nodes with (m1+m2+m3) % 2 == phase % 2 by x += v0koef (b - A x)/a_ii,
 * phase 6 every pressure row by P += p0koef eta_cell (b - A x); A x from the
 * state before the phase (no row of a phase couples to another of the same
 * phase).  Boundary-condition rows are not relaxed. */
void or_colour_phase(const orl_t *L, const double *etaC, const double *etaXY,
		     const double *etaXZ, const double *etaYZ, const int *bondm,
		     const double *b, double *x, int phase, double p0koef, double v0koef)
{
	long N = L->nx * L->ny * L->nz, m1, m2, m3;
	int comp = phase == 6 ? 0 : 1 + phase / 2, col = phase & 1;
	double *nx_ = malloc(sizeof(double) * 4 * N);
	ctx_t C = {L, etaC, etaXY, etaXZ, etaYZ, x};
	memcpy(nx_, x, sizeof(double) * 4 * N);
	for (m3 = 0; m3 < L->nz; m3++)
	for (m1 = 0; m1 < L->nx; m1++)
	for (m2 = 0; m2 < L->ny; m2++) {
		long n = NODE(L, m1, m2, m3), u = n * 4 + comp;
		double ax, d;
		if (!exists(L, m1, m2, m3, comp) || bondm[(L->mgp + n) * 4 + comp]) continue;
		if (comp && ((m1 + m2 + m3) & 1) != col) continue;
		ax = row_ax(&C, m1, m2, m3, comp, &d);
		if (comp) nx_[u] = x[u] + v0koef * (b[u] - ax) / d;
		else nx_[u] = x[u] + p0koef * etaC[n] * (b[u] - ax);
	}
	memcpy(x, nx_, sizeof(double) * 4 * N);
	free(nx_);
}

/* ------------------------------------------------------------------ */
/* Markers                                                            */
/* ------------------------------------------------------------------ */

static long serch(const double *g, long n, double v)
{
	long lo = 0, hi = n - 1, mid;
	do {
		mid = (lo + hi) / 2;
		if (g[mid] > v) hi = mid; else lo = mid;
	} while (hi - lo > 1);
	if (lo > n - 2) lo = n - 2;
	return lo;
}

static double clampx(double v, double hi)
{
	if (v < 0) v = 0;
	if (v > hi) v = hi;
	return v;
}

void or_mk_cells(long n, const float *x, const float *y, const float *z, const double *gx, long nx,
		 const double *gy, long ny, const double *gz, long nz, int *out)
{
	long i;
	for (i = 0; i < n; i++) {
		out[3 * i + 0] = (int)serch(gx, nx, clampx(x[i], gx[nx - 1]));
		out[3 * i + 1] = (int)serch(gy, ny, clampx(y[i], gy[ny - 1]));
		out[3 * i + 2] = (int)serch(gz, nz, clampx(z[i], gz[nz - 1]));
	}
}

/* prcalc1 (mark3mg.c:2917-3001) */
static double or_prcalc(const double *pr, const double *gx, long nx, const double *gy, long ny, const double *gz,
			long nz, double x, double y, double z, long *cell)
{
	long m1, m2, m3, a, b, c, m10, m20, m30;
	double dx, dy, dz, p = 0;
	x = clampx(x, gx[nx - 1]); y = clampx(y, gy[ny - 1]); z = clampx(z, gz[nz - 1]);
	cell[0] = m1 = serch(gx, nx, x);
	cell[1] = m2 = serch(gy, ny, y);
	cell[2] = m3 = serch(gz, nz, z);
	a = m1; if (x > (gx[a] + gx[a + 1]) / 2.0) a += 1; if (a < 1) a = 1; if (a > nx - 2) a = nx - 2;
	b = m2; if (y > (gy[b] + gy[b + 1]) / 2.0) b += 1; if (b < 1) b = 1; if (b > ny - 2) b = ny - 2;
	c = m3; if (z > (gz[c] + gz[c + 1]) / 2.0) c += 1; if (c < 1) c = 1; if (c > nz - 2) c = nz - 2;
	dx = (gx[a + 1] - gx[a - 1]) / 2.0; dy = (gy[b + 1] - gy[b - 1]) / 2.0; dz = (gz[c + 1] - gz[c - 1]) / 2.0;
	for (m30 = c; m30 <= c + 1; m30++) {
		double ddz = ((gz[m30 - 1] + gz[m30]) / 2.0 - z) / dz;
		if (m30 == c) ddz = -ddz;
		ddz = 1.0 - ddz;
		for (m10 = a; m10 <= a + 1; m10++) {
			double ddx = ((gx[m10 - 1] + gx[m10]) / 2.0 - x) / dx;
			if (m10 == a) ddx = -ddx;
			ddx = 1.0 - ddx;
			for (m20 = b; m20 <= b + 1; m20++) {
				double ddy = ((gy[m20 - 1] + gy[m20]) / 2.0 - y) / dy;
				if (m20 == b) ddy = -ddy;
				ddy = 1.0 - ddy;
				p += pr[m30 * nx * ny + m10 * ny + m20] * (ddx * ddy * ddz);
			}
		}
	}
	return p;
}

/* viscalc branches A/B and limits (mark3mg.c:1266-1299, 1626-1681) */
static double or_visc(const or_rock_t *r, double mtk, double mpb, double nubeg, double nuend)
{
	double rt = 8.314 * mtk, nunewt = 0, nueff;
	if (r->nu0) {
		if (r->dh == 0 && r->dv == 0) nunewt = 1.0 / r->nu0;
		else {
			double e1 = (r->dh + r->dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			if (r->dh < 0 && r->dv > 0) { e1 = r->dh * (mtk - r->dv); if (e1 < -150.0) e1 = -150.0; }
			if (r->dh < 0 && r->dv < 0) {
				e1 = (-r->dh) * (1.0 - (mtk - (-r->dv)) / (-r->dv)) / 8.314 / (-r->dv);
				if (e1 > 150.0) e1 = 150.0;
			}
			nunewt = 1.0 / (r->nu0 * exp(e1));
		}
	}
	nueff = nunewt;
	if (nueff <= 0) nueff = 1.0 / r->n1;
	nueff = 1.0 / nueff;
	if (nueff < r->n0) nueff = r->n0;
	if (nueff > r->n1) nueff = r->n1;
	if (nueff < nubeg) nueff = nubeg;
	if (nueff > nuend) nueff = nuend;
	return nueff;
}

void or_ronurecalc(long n, const float *x, const float *y, const float *z, const float *tk, float *eta,
		   const signed char *type, const double *gx, long nx, const double *gy, long ny, const double *gz,
		   long nz, const double *pr, const or_rock_t *rocks, double nubeg, double nuend, double nukoef,
		   int densimod, double timestep, double vdeep, double nudeep, float *fld)
{
	long N = nx * ny * nz, i, q;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	float *acc = calloc((size_t)(11 * N), sizeof(float));
	double *accd = calloc((size_t)(11 * N), sizeof(double));
	double *w = calloc((size_t)(6 * N), sizeof(double)); /* w1 w0 wxx wxy wxz wyz */
	for (i = 0; i < n; i++) {
		long cell[3], m1, m2, m3, m10, m20, m30;
		double mpb, mtk, mnu, mro, mbb = 0, mcp, mkt, mht, dx, dy, dz;
		const or_rock_t *r;
		int mt = type[i];
		if (!(x[i] > 0 && y[i] > 0 && z[i] > 0 && x[i] < xs && y[i] < ys && z[i] < zs && mt < 50 && tk[i] > 0))
			continue;
		r = &rocks[mt];
		mpb = or_prcalc(pr, gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], cell) * 1e-5;
		mtk = tk[i];
		mnu = eta[i];
		if (mnu <= 0 || timestep <= 0) {
			mnu = or_visc(r, mtk, mpb, nubeg, nuend);
			eta[i] = (float)mnu;
		}
		if (densimod == 0) mro = r->ro;
		else {
			mro = r->ro * (1.0 - r->bro * (mtk - 298.15)) * (1.0 + r->aro * (mpb - 1.0) * 1e-3);
			mbb = r->bro / (1.0 - r->bro * (mtk - 298.15));
		}
		mcp = r->cp;
		mkt = (r->kt + r->kf / (mtk + 77.0)) * exp(r->kp * mpb);
		mht = r->ht;
		if (y[i] > vdeep && mnu > nudeep) mnu = nudeep;
		if (y[i] < 5e+4 && x[i] > xs - 1e+4 && z[i] > 1.5e+5 && z[i] < 8.7e+5 && mnu > 1e+19) mnu = 1e+19;
		m10 = cell[0]; m20 = cell[1]; m30 = cell[2];
		dx = gx[m10 + 1] - gx[m10]; dy = gy[m20 + 1] - gy[m20]; dz = gz[m30 + 1] - gz[m30];
		for (m3 = m30; m3 <= m30 + 1; m3++) {
			double ddz = 1.0 - fabs((gz[m3] - z[i]) / dz);
			for (m1 = m10; m1 <= m10 + 1; m1++) {
				double ddx = 1.0 - fabs((gx[m1] - x[i]) / dx);
				for (m2 = m20; m2 <= m20 + 1; m2++) {
					double ddy = 1.0 - fabs((gy[m2] - y[i]) / dy);
					double swt = ddx * ddy * ddz, s1;
					long m4 = m3 * nx * ny + m1 * ny + m2;
					if (ddx > nukoef && ddy > nukoef && ddz > nukoef) {
						or_acc_add(acc, accd, m4, mnu * swt);
						w[m4] += swt;
					}
					if (m1 > m10 && m2 > m20 && m3 > m30) {
						s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (0.5 - fabs(ddy - 0.5)) * (0.5 - fabs(ddz - 0.5));
						or_acc_add(acc, accd, 7 * N + m4, mnu * s1);
						w[2 * N + m4] += s1;
					}
					if (m3 == m30 && ddx > 0.5 && ddy > 0.5) {
						s1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - fabs(ddz - 0.5));
						or_acc_add(acc, accd, 8 * N + m4, mnu * s1);
						w[3 * N + m4] += s1;
					}
					if (m2 == m20 && ddx > 0.5 && ddz > 0.5) {
						s1 = 8.0 * (ddx - 0.5) * (0.5 - fabs(ddy - 0.5)) * (ddz - 0.5);
						or_acc_add(acc, accd, 9 * N + m4, mnu * s1);
						w[4 * N + m4] += s1;
					}
					if (m1 == m10 && ddy > 0.5 && ddz > 0.5) {
						s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
						or_acc_add(acc, accd, 10 * N + m4, mnu * s1);
						w[5 * N + m4] += s1;
					}
					if (swt > 0) {
						or_acc_add(acc, accd, 1 * N + m4, mro * swt);
						or_acc_add(acc, accd, 2 * N + m4, mcp * mro * swt);
						or_acc_add(acc, accd, 3 * N + m4, mkt * swt);
						or_acc_add(acc, accd, 4 * N + m4, mht * swt);
						or_acc_add(acc, accd, 5 * N + m4, mbb * swt);
						or_acc_add(acc, accd, 6 * N + m4, mtk * swt);
						w[N + m4] += swt;
					}
				}
			}
		}
	}
	for (i = 0; i < N; i++) {
		if (w[i] > 0.01) fld[i] = (float)(or_acc_get(acc, accd, i) / w[i]);
		for (q = 0; q < 4; q++)
			if (w[(2 + q) * N + i] > 0.01) fld[(7 + q) * N + i] = (float)(or_acc_get(acc, accd, (7 + q) * N + i) / w[(2 + q) * N + i]);
		if (w[N + i] > 0.01) {
			fld[N + i] = (float)(or_acc_get(acc, accd, N + i) / w[N + i]);
			fld[2 * N + i] = (float)(or_acc_get(acc, accd, 2 * N + i) / w[N + i] / fld[N + i]);
			for (q = 3; q <= 6; q++) fld[q * N + i] = (float)(or_acc_get(acc, accd, q * N + i) / w[N + i]);
		}
		if (fld[i] < nubeg) fld[i] = (float)nubeg;
		if (fld[i] > nuend) fld[i] = (float)nuend;
	}
	free(acc);
	free(accd);
	free(w);
}

void or_vxyz(const float *vx, const float *vy, const float *vz, const double *gx, long nx, const double *gy,
	     long ny, const double *gz, long nz, double x, double y, double z, double *v)
{
	long m1, m2, m3, comp;
	x = clampx(x, gx[nx - 1]); y = clampx(y, gy[ny - 1]); z = clampx(z, gz[nz - 1]);
	m1 = serch(gx, nx, x); m2 = serch(gy, ny, y); m3 = serch(gz, nz, z);
	for (comp = 0; comp < 3; comp++) {
		long a = m1, b = m2, c = m3, m10, m20, m30;
		double dx, dy, dz, acc = 0;
		const float *f = comp == 0 ? vx : comp == 1 ? vy : vz;
		if (comp == 0) { if (a > nx - 2) a = nx - 2; }
		else { if (x < (gx[a] + gx[a + 1]) / 2.0) a -= 1; if (a < 0) a = 0; if (a > nx - 3) a = nx - 3; }
		if (comp == 1) { if (b > ny - 2) b = ny - 2; }
		else { if (y < (gy[b] + gy[b + 1]) / 2.0) b -= 1; if (b < 0) b = 0; if (b > ny - 3) b = ny - 3; }
		if (comp == 2) { if (c > nz - 2) c = nz - 2; }
		else { if (z < (gz[c] + gz[c + 1]) / 2.0) c -= 1; if (c < 0) c = 0; if (c > nz - 3) c = nz - 3; }
		dx = comp == 0 ? gx[a + 1] - gx[a] : (gx[a + 2] - gx[a]) / 2.0;
		dy = comp == 1 ? gy[b + 1] - gy[b] : (gy[b + 2] - gy[b]) / 2.0;
		dz = comp == 2 ? gz[c + 1] - gz[c] : (gz[c + 2] - gz[c]) / 2.0;
		for (m30 = c; m30 <= c + 1; m30++) {
			double ddz = comp == 2 ? (gz[m30] - z) / dz : ((gz[m30] + gz[m30 + 1]) / 2.0 - z) / dz;
			if (m30 == c) ddz = -ddz;
			ddz = 1.0 - ddz;
			for (m10 = a; m10 <= a + 1; m10++) {
				double ddx = comp == 0 ? (gx[m10] - x) / dx : ((gx[m10] + gx[m10 + 1]) / 2.0 - x) / dx;
				if (m10 == a) ddx = -ddx;
				ddx = 1.0 - ddx;
				for (m20 = b; m20 <= b + 1; m20++) {
					double ddy = comp == 1 ? (gy[m20] - y) / dy : ((gy[m20] + gy[m20 + 1]) / 2.0 - y) / dy;
					if (m20 == b) ddy = -ddy;
					ddy = 1.0 - ddy;
					acc += f[m30 * nx * ny + m10 * ny + m20] * (ddx * ddy * ddz);
				}
			}
		}
		v[comp] = acc;
	}
}

void or_rk4(long n, float *x, float *y, float *z, float *tk, float *eta, signed char *type,
	    const float *vx, const float *vy, const float *vz, const double *gx, long nx, const double *gy,
	    long ny, const double *gz, long nz, double dt, double zdeep, double vdeep, double tdeep)
{
	long i;
	int q;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	for (i = 0; i < n; i++) {
		double v1[3], v2[3], v3[3], v4[3], v[3], px = x[i], py = y[i], pz = z[i];
		int mt = type[i];
		if (!((x[i] >= 0 && y[i] >= 0 && z[i] >= 0 && x[i] <= xs && y[i] <= ys && z[i] <= zs) || mt < 50 || mt >= 100))
			continue;
		or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px, py, pz, v1);
		or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v1[0] * dt / 2.0, py + v1[1] * dt / 2.0, pz + v1[2] * dt / 2.0, v2);
		or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v2[0] * dt / 2.0, py + v2[1] * dt / 2.0, pz + v2[2] * dt / 2.0, v3);
		or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v3[0] * dt, py + v3[1] * dt, pz + v3[2] * dt, v4);
		for (q = 0; q < 3; q++) v[q] = (v1[q] + 2.0 * v2[q] + 2.0 * v3[q] + v4[q]) / 6.0;
		if (mt < 100) {
			double yold = y[i];
			if (y[i] > zdeep && v[1] < 0 && tk[i] < tdeep && type[i] == 10) tk[i] = (float)tdeep;
			if (y[i] > zdeep && v[1] < 0 && type[i] != 10) type[i] = 10;
			if (y[i] > zdeep && type[i] != 10 && type[i] != 9) type[i] = 9;
			x[i] += dt * v[0] - 0.0;
			y[i] += dt * v[1];
			z[i] += dt * v[2];
			if ((yold > vdeep && y[i] < vdeep) || (yold > zdeep && y[i] < zdeep)) {
				if (tk[i] < tdeep) tk[i] = (float)tdeep;
				type[i] = 10;
			}
			if (x[i] < 0 || y[i] < 0 || z[i] < 0 || x[i] > xs || y[i] > ys || z[i] > zs) {
				tk[i] = 0;
				eta[i] = 0;
			}
		}
	}
}

void or_tkrecalc(long nodenum, const int *bondm, const float *bondv, const int *bondn, float *tk)
{
	long m;
	float *tk1 = malloc(sizeof(float) * (size_t)nodenum);
	memcpy(tk1, tk, sizeof(float) * (size_t)nodenum);
	for (m = 0; m < nodenum; m++) {
		int rec = bondm[8 * nodenum + m];
		if (rec > 0) {
			tk[m] = bondv[2 * rec];
			if (bondn[rec]) tk[m] += bondv[2 * rec + 1] * tk1[bondn[rec] - 1];
		}
	}
	free(tk1);
}

/* ---- heat (test infrastructure) ----------------------------------------- */

/* Row of level-0 node (m1,m2,m3) of the implicit heat system, in the order
 * the reference stores it: heaterr(ynerr=1) heat3mg.c:160-225 for interior
 * nodes (diagonal first, then x-, x+, y-, y+, z-, z+), tbonderr(ynerr=0)
 * heat3mg.c:818-841 for boundary nodes (1 at the node, -koef at the
 * partner), with the bottom-line plume override of titerate
 * heat3mg.c:444-455.  Returns the entry count. */
static int heat_row(const or_heat_t *H, long m1, long m2, long m3, int *col, double *val, double *rhs)
{
	long nx = H->nx, ny = H->ny, nxy = nx * ny;
	long v3 = m3 * nxy + m1 * ny + m2;
	int rec = H->bondm_t[v3];
	if (rec > 0) {
		double c = H->bondv[2 * rec];
		int n = 1;
		col[0] = (int)v3;
		val[0] = 1.0;
		if (H->bondn[rec]) {
			col[1] = H->bondn[rec] - 1;
			val[1] = -(double)H->bondv[2 * rec + 1];
			n = 2;
		}
		if (H->timesum > H->timebeg && m2 == ny - 1) {
			if (H->timesum >= H->timeend) c = H->tempend;
			else c = c + (H->tempend - c) / (H->timeend - H->timebeg) * (H->timesum - H->timebeg);
		}
		*rhs = c;
		return n;
	}
	const double *gx = H->gx, *gy = H->gy, *gz = H->gz;
	const float *kt = H->kt;
	double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
	double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
	double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
	long v[7] = {v3 - nxy, v3 - ny, v3 - 1, v3, v3 + 1, v3 + ny, v3 + nxy};
	double rocp = (float)(H->ro[v3] * H->cp[v3]);
	double iv;
	*rhs = (H->src ? H->src[v3] : (double)H->ht[v3]) + H->tk0[v3] / H->dt * rocp;
	val[0] = rocp / H->dt;
	col[0] = (int)v3;
	iv = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[m1] - gx[m1 - 1]);
	val[1] = -iv; col[1] = (int)v[1]; val[0] += iv;
	iv = (float)(kt[v[3]] + kt[v[5]]) / 2.0 * xkf / (gx[m1 + 1] - gx[m1]);
	val[0] += iv; val[2] = -iv; col[2] = (int)v[5];
	iv = (float)(kt[v[2]] + kt[v[3]]) / 2.0 * ykf / (gy[m2] - gy[m2 - 1]);
	val[3] = -iv; col[3] = (int)v[2]; val[0] += iv;
	iv = (float)(kt[v[3]] + kt[v[4]]) / 2.0 * ykf / (gy[m2 + 1] - gy[m2]);
	val[0] += iv; val[4] = -iv; col[4] = (int)v[4];
	iv = (float)(kt[v[0]] + kt[v[3]]) / 2.0 * zkf / (gz[m3] - gz[m3 - 1]);
	val[5] = -iv; col[5] = (int)v[0]; val[0] += iv;
	iv = (float)(kt[v[3]] + kt[v[6]]) / 2.0 * zkf / (gz[m3 + 1] - gz[m3]);
	val[0] += iv; val[6] = -iv; col[6] = (int)v[6];
	return 7;
}

void or_heat_rows(const or_heat_t *H, int *cnt, int *cols, double *vals, double *rhs)
{
	long m1, m2, m3;
	for (m3 = 0; m3 < H->nz; m3++)
		for (m1 = 0; m1 < H->nx; m1++)
			for (m2 = 0; m2 < H->ny; m2++) {
				long v = m3 * H->nx * H->ny + m1 * H->ny + m2;
				cnt[v] = heat_row(H, m1, m2, m3, cols + 7 * v, vals + 7 * v, rhs + v);
			}
}

/* Step limit of titerate (heat3mg.c:327-368): largest |H/(ro cp)| over the
 * non-boundary nodes (heaterr ynerr=0 without adiabatic/shear terms). */
double or_heat_maxrate(const or_heat_t *H)
{
	long v, N = H->nx * H->ny * H->nz;
	double mx = 0;
	for (v = 0; v < N; v++)
		if (H->bondm_t[v] == 0) {
			double rocp = (float)(H->ro[v] * H->cp[v]);
			double r = fabs((H->src ? H->src[v] : (double)H->ht[v]) / rocp);
			if (r > mx) mx = r;
		}
	return mx;
}

/* Convergence norms of titerate (heat3mg.c:512-560): heatsum = RMS over
 * non-boundary nodes of heaterr(ynerr=2)*dt with tk stored as float,
 * bondsum = RMS of the boundary-row errors (tbonderr ynerr=1 on the
 * double iterate).  out[0] = heatsum, out[1] = bondsum. */
void or_heat_norms(const or_heat_t *H, const double *x, double *out)
{
	long nx = H->nx, ny = H->ny, nz = H->nz, nxy = nx * ny, N = nxy * nz;
	long m1, m2, m3;
	double hs = 0, hn = 0, bs = 0, bn = 0;
	float *tk = malloc(sizeof(float) * (size_t)N);
	for (m1 = 0; m1 < N; m1++) tk[m1] = (float)x[m1];
	const double *gx = H->gx, *gy = H->gy, *gz = H->gz;
	const float *kt = H->kt;
	for (m3 = 0; m3 < nz; m3++)
		for (m1 = 0; m1 < nx; m1++)
			for (m2 = 0; m2 < ny; m2++) {
				long v3 = m3 * nxy + m1 * ny + m2;
				int rec = H->bondm_t[v3];
				if (rec > 0) {
					double e = x[v3] - H->bondv[2 * rec];
					if (H->bondn[rec]) e -= H->bondv[2 * rec + 1] * x[H->bondn[rec] - 1];
					bs += e * e;
					bn += 1;
					continue;
				}
				long v[7] = {v3 - nxy, v3 - ny, v3 - 1, v3, v3 + 1, v3 + ny, v3 + nxy};
				double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
				double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
				double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
				double rocp = (float)(H->ro[v3] * H->cp[v3]);
				double e = (double)(tk[v3] - H->tk0[v3]) / H->dt, iv;
				iv = (float)(kt[v[1]] + kt[v[3]]) / 2.0 * xkf / (gx[m1] - gx[m1 - 1]) / rocp;
				e += (double)(tk[v3] - tk[v[1]]) * iv;
				iv = (float)(kt[v[3]] + kt[v[5]]) / 2.0 * xkf / (gx[m1 + 1] - gx[m1]) / rocp;
				e -= (double)(tk[v[5]] - tk[v3]) * iv;
				iv = (float)(kt[v[2]] + kt[v[3]]) / 2.0 * ykf / (gy[m2] - gy[m2 - 1]) / rocp;
				e += (double)(tk[v3] - tk[v[2]]) * iv;
				iv = (float)(kt[v[3]] + kt[v[4]]) / 2.0 * ykf / (gy[m2 + 1] - gy[m2]) / rocp;
				e -= (double)(tk[v[4]] - tk[v3]) * iv;
				iv = (float)(kt[v[0]] + kt[v[3]]) / 2.0 * zkf / (gz[m3] - gz[m3 - 1]) / rocp;
				e += (double)(tk[v3] - tk[v[0]]) * iv;
				iv = (float)(kt[v[3]] + kt[v[6]]) / 2.0 * zkf / (gz[m3 + 1] - gz[m3]) / rocp;
				e -= (double)(tk[v[6]] - tk[v3]) * iv;
				e -= (H->src ? H->src[v3] : (double)H->ht[v3]) / rocp; /* tk1 */
				e *= H->dt;
				hs += e * e;
				hn += 1;
			}
	free(tk);
	out[0] = sqrt(hs / hn);
	out[1] = bn > 0 ? sqrt(bs / bn) : 0;
}

/* tkcalc1 (mark3mg.c:2844-2910): trilinear weights of the 8 nodes of the
 * cell holding (x,y,z); returns the cell and spacings. */
static void tk_weights(const double *gx, long nx, const double *gy, long ny, const double *gz, long nz, double x,
		       double y, double z, long *cell, double *d, long *wn, double *wi)
{
	long m10 = serch(gx, nx, x), m20 = serch(gy, ny, y), m30 = serch(gz, nz, z);
	long a, b, c;
	int q = 0;
	cell[0] = m10; cell[1] = m20; cell[2] = m30;
	d[0] = gx[m10 + 1] - gx[m10];
	d[1] = gy[m20 + 1] - gy[m20];
	d[2] = gz[m30 + 1] - gz[m30];
	for (c = m30; c <= m30 + 1; c++) {
		double ddz = 1.0 - fabs((gz[c] - z) / d[2]);
		for (a = m10; a <= m10 + 1; a++) {
			double ddx = 1.0 - fabs((gx[a] - x) / d[0]);
			for (b = m20; b <= m20 + 1; b++) {
				double ddy = 1.0 - fabs((gy[b] - y) / d[1]);
				wi[q] = ddx * ddy * ddz;
				wn[q] = c * nx * ny + a * ny + b;
				q++;
			}
		}
	}
}

/* Marker temperature update after the heat solve (titerate,
 * heat3mg.c:592-782): markers of type 50..99 inside the grid take the
 * interpolated new temperature; for markers of type < 50 strictly inside:
 * new markers (markk <= 0) take the interpolated start temperature (deep
 * floor tdeep below zdeep); with heatdif > 0 the subgrid diffusion relaxes
 * the marker towards the grid, its change is averaged back onto the nodes
 * in marker order and removed again (antidiffusion); finally the marker
 * takes the interpolated temperature change.  The marker loops run over
 * n markers (the reference visits one index past marknum, which the caller
 * keeps outside the grid). */
void or_mk_tk_update(long n, const float *x, const float *y, const float *z, const signed char *type, float *markk,
		     const double *gx, long nx, const double *gy, long ny, const double *gz, long nz, const float *tk,
		     const float *tk2, const float *kt, const float *ro, const float *cp, double heatdif, double dt,
		     double zdeep, double tdeep)
{
	long N = nx * ny * nz, i;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	double *s0 = calloc((size_t)N, sizeof(double)), *s1 = calloc((size_t)N, sizeof(double));
	/* fluid markers (50 <= type < 100) take the new grid temperature */
	for (i = 0; i < n; i++) {
		if (!(type[i] >= 50 && type[i] < 100 && x[i] > 0 && y[i] > 0 && z[i] > 0 && (double)x[i] < xs &&
		      (double)y[i] < ys && (double)z[i] < zs))
			continue;
		long cell[3], wn[8];
		double d[3], wi[8], e2 = 0;
		int q;
		tk_weights(gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], cell, d, wn, wi);
		for (q = 0; q < 8; q++) e2 += wi[q] * tk[wn[q]];
		markk[i] = (float)e2;
	}
	for (i = 0; i < n; i++) {
		if (!(type[i] < 50 && x[i] > 0 && y[i] > 0 && z[i] > 0 && (double)x[i] < xs && (double)y[i] < ys &&
		      (double)z[i] < zs))
			continue;
		long cell[3], wn[8];
		double d[3], wi[8], e2 = 0, e3 = 0;
		int q;
		tk_weights(gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], cell, d, wn, wi);
		for (q = 0; q < 8; q++) {
			e2 += wi[q] * tk[wn[q]];
			e3 += wi[q] * tk2[wn[q]];
		}
		if (markk[i] <= 0) {
			markk[i] = (float)e3;
			if (y[i] > zdeep && markk[i] < tdeep) markk[i] = (float)tdeep;
		}
		if (heatdif) {
			double mkt = 0, rocp = 0, mtk = tk[wn[0]], dtk, dtk1;
			for (q = 0; q < 8; q++) {
				mkt += kt[wn[q]] * wi[q];
				rocp += ro[wn[q]] * cp[wn[q]] * wi[q];
				if (tk[wn[q]] > mtk) mtk = tk[wn[q]];
			}
			if ((cell[0] <= 0 || cell[0] >= nx - 2 || cell[1] <= 0 || cell[1] >= ny - 2 || cell[2] <= 0 ||
			     cell[2] >= nz - 2) &&
			    (markk[i] + e2 - e3) > mtk)
				markk[i] = (float)(mtk - e2 + e3);
			dtk = e3 - (double)markk[i];
			dtk1 = -heatdif * mkt / rocp * 2.0 * (1.0 / d[0] / d[0] + 1.0 / d[1] / d[1] + 1.0 / d[2] / d[2]) * dt;
			if (dtk1 < -150.0) dtk1 = -150.0;
			dtk1 = dtk * (1.0 - exp(dtk1));
			markk[i] += (float)dtk1;
			for (q = 0; q < 8; q++) {
				s1[wn[q]] += dtk1 * wi[q];
				s0[wn[q]] += wi[q];
			}
		}
		markk[i] += (float)(e2 - e3);
	}
	if (heatdif) {
		for (i = 0; i < N; i++)
			if (s0[i]) s1[i] /= s0[i];
		for (i = 0; i < n; i++) {
			if (!(type[i] < 50 && x[i] > 0 && y[i] > 0 && z[i] > 0 && (double)x[i] < xs && (double)y[i] < ys &&
			      (double)z[i] < zs))
				continue;
			long cell[3], wn[8];
			double d[3], wi[8], a = 0;
			int q;
			tk_weights(gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], cell, d, wn, wi);
			for (q = 0; q < 8; q++) a -= s1[wn[q]] * wi[q];
			markk[i] += (float)a;
		}
	}
	free(s0);
	free(s1);
}

/* ---- native norms (test infrastructure) -------------------------------- */

/* 48-weight viscosity of a momentum row (x/y/zstokserr ynerr=1,
 * move3mg.c:770, 966, 1165): the reference sums the first four node values
 * in float, the ten-fold group in float, the rest in double. */
static double nueff48(const float *nu, long a0, long a1, long a2, long a3, long b0, long b1, long b2, long b3,
		      long c0, long c1, long c2, long c3, double nubeg, double nuend)
{
	float s1 = nu[a0] + nu[a1] + nu[a2] + nu[a3];
	float s2 = nu[b0] + nu[b1] + nu[b2] + nu[b3];
	double e = ((double)s1 + 10.0 * (double)s2 + nu[c0] + nu[c1] + nu[c2] + nu[c3]) / 48.0;
	if (e > nuend) e = nuend;
	if (e < nubeg) e = nubeg;
	return e;
}

void or_err_rows(const orl_t *L, const float *nu, const float *ro, const double *g, double nubeg, double nuend,
		 const float *st12, const double *pr, const float *vx, const float *vy, const float *vz,
		 const int *bondm, const float *bondv, const int *bondn, const double *sol, double *rows)
{
	long nx = L->nx, ny = L->ny, nz = L->nz, N = nx * ny * nz, sx = ny, sz = nx * ny;
	const double *gx = L->gx, *gy = L->gy, *gz = L->gz;
	const float *sxx = st12 + 6 * N, *syy = st12 + 7 * N, *szz = st12 + 8 * N;
	const float *sxy = st12 + 9 * N, *sxz = st12 + 10 * N, *syz = st12 + 11 * N;
	long m1, m2, m3, i;
	for (i = 0; i < 5 * 4 * N * 2; i++) rows[i] = NAN;
#define ROW(k, u, a, b) do { rows[(((long)(k) * 4 * N) + (u)) * 2] = (a); rows[(((long)(k) * 4 * N) + (u)) * 2 + 1] = (b); } while (0)
	for (m3 = 0; m3 < nz; m3++)
	for (m1 = 0; m1 < nx; m1++)
	for (m2 = 0; m2 < ny; m2++) {
		long n = m3 * sz + m1 * sx + m2;
		if (m1 && m2 && m3) {
			/* continerr ynerr=1 (move3mg.c:1342-1348) */
			double xkf = 1.0 / (gx[m1] - gx[m1 - 1]), ykf = 1.0 / (gy[m2] - gy[m2 - 1]),
			       zkf = 1.0 / (gz[m3] - gz[m3 - 1]);
			long v0 = n - sz - sx - 1;
			double lc = (vx[v0 + sx] - vx[v0]) * xkf + (vy[v0 + 1] - vy[v0]) * ykf + (vz[v0 + sz] - vz[v0]) * zkf;
			double skf = (1.0 / xkf + 1.0 / ykf + 1.0 / zkf) / 3.0;
			ROW(0, 4 * n, (lc - 0) * skf, 0.0);
		}
		for (int c = 1; c <= 3; c++) {
			int ok = c == 1 ? (m2 < ny - 1 && m3 < nz - 1) : c == 2 ? (m1 < nx - 1 && m3 < nz - 1)
									: (m1 < nx - 1 && m2 < ny - 1);
			if (!ok) continue;
			long u = 4 * n + c;
			int rec = bondm[u];
			if (rec) {
				double lb = sol[u] - bondv[2 * rec];
				if (bondn[rec]) lb -= bondv[2 * rec + 1] * sol[bondn[rec] - 1];
				ROW(4, u, lb, 0.0);
				continue;
			}
			long v[8];
			v[0] = n; v[1] = n + 1; v[2] = n + sx; v[3] = n + sx + 1;
			v[4] = n + sz; v[5] = n + sz + 1; v[6] = n + sz + sx; v[7] = n + sz + sx + 1;
			double xkf, ykf, zkf, right, left, ne;
			if (c == 1) {
				xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]); ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
				zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
				right = -g[0] * (ro[v[0]] + ro[v[1]] + ro[v[4]] + ro[v[5]]) / 4.0;
				ne = nueff48(nu, v[0] - sx, v[1] - sx, v[4] - sx, v[5] - sx, v[0], v[1], v[4], v[5], v[2], v[3],
					     v[6], v[7], nubeg, nuend);
				left = (sxx[v[7]] - sxx[v[5]]) * xkf;
				left += (sxy[v[1]] - sxy[v[0]]) * ykf;
				left += (sxz[v[4]] - sxz[v[0]]) * zkf;
				left -= (pr[v[7]] - pr[v[5]]) * xkf;
			} else if (c == 2) {
				xkf = 1.0 / (gx[m1 + 1] - gx[m1]); ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
				zkf = 1.0 / (gz[m3 + 1] - gz[m3]);
				right = -g[1] * (ro[v[0]] + ro[v[2]] + ro[v[4]] + ro[v[6]]) / 4.0;
				ne = nueff48(nu, v[0] - 1, v[2] - 1, v[4] - 1, v[6] - 1, v[0], v[2], v[4], v[6], v[1], v[3],
					     v[5], v[7], nubeg, nuend);
				left = (syy[v[7]] - syy[v[6]]) * ykf;
				left += (sxy[v[2]] - sxy[v[0]]) * xkf;
				left += (syz[v[4]] - syz[v[0]]) * zkf;
				left -= (pr[v[7]] - pr[v[6]]) * ykf;
			} else {
				xkf = 1.0 / (gx[m1 + 1] - gx[m1]); ykf = 1.0 / (gy[m2 + 1] - gy[m2]);
				zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
				right = -g[2] * (ro[v[0]] + ro[v[1]] + ro[v[2]] + ro[v[3]]) / 4.0;
				ne = nueff48(nu, v[0] - sz, v[1] - sz, v[2] - sz, v[3] - sz, v[0], v[1], v[2], v[3], v[4], v[5],
					     v[6], v[7], nubeg, nuend);
				left = (szz[v[7]] - szz[v[3]]) * zkf;
				left += (sxz[v[2]] - sxz[v[0]]) * xkf;
				left += (syz[v[1]] - syz[v[0]]) * ykf;
				left -= (pr[v[7]] - pr[v[3]]) * zkf;
			}
			double skf1 = 1.0 / ne * (1.0 / xkf / xkf + 1.0 / ykf / ykf + 1.0 / zkf / zkf) / 3.0;
			left = left - right;
			ROW(c, u, left, left * skf1);
		}
	}
#undef ROW
}

/* heat3mg.c:54-160.  Float products stay float where the reference
 * multiplies two float arrays (et*tk0, ro*cp, sigma*eps). */
void or_heat_source(long nx, long ny, long nz, const double *gx, const double *gy, const double *gz,
		    const float *ro, const float *cp, const float *ht, const float *tk0, const float *et,
		    const float *vx, const float *vy, const float *vz, const double *pr, const float *const *s6,
		    const float *const *e6, const int *bondm_t, int adiabyn, int frictyn, const double *gk,
		    double ysize, double stp100, float *src, double *rate)
{
	long nxy = nx * ny, m1, m2, m3;
	for (m3 = 0; m3 < nz; m3++)
	for (m1 = 0; m1 < nx; m1++)
	for (m2 = 0; m2 < ny; m2++) {
		long v3 = m3 * nxy + m1 * ny + m2;
		src[v3] = 0;
		rate[v3] = 0;
		if (bondm_t[v3] != 0) continue;
		double xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
		double ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
		double zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
		double rocpkf = (float)(ro[v3] * cp[v3]);
		double ival = 0, x = 0, y = 0, z = 0, v[3] = {0, 0, 0};
		/* P cell centres around the node, corner order y fastest, then x,
		 * then z: c[0] = v3, c[1] = +y, c[2] = +x, ..., c[7] = +x+y+z */
		long c[8];
		int q;
		for (q = 0; q < 8; q++) c[q] = v3 + (q & 1) + ((q >> 1) & 1) * ny + ((q >> 2) & 1) * nxy;
		ival += ht[v3];
		if (adiabyn == 2 || frictyn) {
			x = (gx[m1] - gx[m1 - 1]) / 2.0 * xkf;
			y = (gy[m2] - gy[m2 - 1]) / 2.0 * ykf;
			z = (gz[m3] - gz[m3 - 1]) / 2.0 * zkf;
		}
		if (adiabyn) or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, gx[m1], gy[m2], gz[m3], v);
		int act = ro[v3] > 2e+3 && gy[m2] < ysize - stp100;
		if (adiabyn == 2 && act) {
			/* trilinear P gradient between the cell centres */
			double dpdx = xkf * ((1.0 - y) * (1.0 - z) * (pr[c[2]] - pr[c[0]]) + y * (1.0 - z) * (pr[c[3]] - pr[c[1]]) +
					     (1.0 - y) * z * (pr[c[6]] - pr[c[4]]) + y * z * (pr[c[7]] - pr[c[5]]));
			double dpdy = ykf * ((1.0 - x) * (1.0 - z) * (pr[c[1]] - pr[c[0]]) + x * (1.0 - z) * (pr[c[3]] - pr[c[2]]) +
					     (1.0 - x) * z * (pr[c[5]] - pr[c[4]]) + x * z * (pr[c[7]] - pr[c[6]]));
			double dpdz = zkf * ((1.0 - x) * (1.0 - y) * (pr[c[4]] - pr[c[0]]) + x * (1.0 - y) * (pr[c[6]] - pr[c[2]]) +
					     (1.0 - x) * y * (pr[c[5]] - pr[c[1]]) + x * y * (pr[c[7]] - pr[c[3]]));
			float at = et[v3] * tk0[v3];
			ival += at * (v[0] * dpdx + v[1] * dpdy + v[2] * dpdz);
		}
		if (adiabyn == 1 && act) {
			float atr = et[v3] * tk0[v3] * ro[v3];
			ival += atr * (v[0] * gk[0] + v[1] * gk[1] + v[2] * gk[2]);
		}
		if (frictyn && act) {
			/* normal terms: trilinear over the eight centres */
			double wx[2] = {1.0 - x, x}, wy[2] = {1.0 - y, y}, wz[2] = {1.0 - z, z};
			double e = 0;
			for (q = 0; q < 8; q++) {
				long u = c[q];
				float se = s6[0][u] * e6[0][u] + s6[1][u] * e6[1][u] + s6[2][u] * e6[2][u];
				e += wx[(q >> 1) & 1] * wy[q & 1] * wz[(q >> 2) & 1] * se;
			}
			/* shear terms: the two edge nodes along the third axis */
			e += 2.0 * ((1.0 - z) * (s6[3][v3 - nxy] * e6[3][v3 - nxy]) + z * (s6[3][v3] * e6[3][v3]));
			e += 2.0 * ((1.0 - y) * (s6[4][v3 - 1] * e6[4][v3 - 1]) + y * (s6[4][v3] * e6[4][v3]));
			e += 2.0 * ((1.0 - x) * (s6[5][v3 - ny] * e6[5][v3 - ny]) + x * (s6[5][v3] * e6[5][v3]));
			ival += e;
		}
		src[v3] = (float)ival;
		rate[v3] = ival / rocpkf;
	}
}

/* One axis of an epscalc1 interpolation block: the interpolation points are
 * cell centres (kind 0, base cell + 1 past the mid-line, base in
 * [1, n-2]), grid lines (kind 1, base in [1, n-3]) or mid-lines (kind 2,
 * base cell - 1 below the mid-line, base in [0, n-3]); w[0..1] are the
 * clamped weights of base and base+1 (mark3mg.c:2575-2600, 2646-2662). */
static long eps_axis(const double *g, long n, long cell, double x, int kind, double *w)
{
	long b = cell;
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

void or_epscalc(const float *const *st12, const double *pr, const double *gx, long nx, const double *gy, long ny,
		const double *gz, long nz, double x, double y, double z, double *out)
{
	/* axis kinds of the four blocks: centres, xy edges, xz edges, yz edges */
	static const int K[4][3] = {{0, 0, 0}, {1, 1, 2}, {1, 2, 1}, {2, 1, 1}};
	const float *e[6], *s[6];
	long nxy = nx * ny, q, blk;
	for (q = 0; q < 6; q++) { e[q] = st12[q]; s[q] = st12[6 + q]; }
	for (q = 0; q < 14; q++) out[q] = 0;
	x = clampx(x, gx[nx - 1]); y = clampx(y, gy[ny - 1]); z = clampx(z, gz[nz - 1]);
	long m1 = serch(gx, nx, x), m2 = serch(gy, ny, y), m3 = serch(gz, nz, z);
	for (blk = 0; blk < 4; blk++) {
		double wx[2], wy[2], wz[2];
		long a = eps_axis(gx, nx, m1, x, K[blk][0], wx);
		long b = eps_axis(gy, ny, m2, y, K[blk][1], wy);
		long c = eps_axis(gz, nz, m3, z, K[blk][2], wz);
		long i, j, k;
		for (k = 0; k < 2; k++)
		for (i = 0; i < 2; i++)
		for (j = 0; j < 2; j++) {
			long u = (c + k) * nxy + (a + i) * ny + (b + j);
			double swt = wx[i] * wy[j] * wz[k];
			if (blk == 0) {
				out[0] += e[0][u] * swt;
				out[1] += e[1][u] * swt;
				out[2] += e[2][u] * swt;
				out[7] += s[0][u] * swt;
				out[8] += s[1][u] * swt;
				out[9] += s[2][u] * swt;
				out[13] += (s[0][u] * e[0][u] + s[1][u] * e[1][u] + s[2][u] * e[2][u]) * swt;
				out[6] += pr[u] * swt;
			} else {
				/* blk 1: xy, 2: xz, 3: yz */
				out[2 + blk] += e[2 + blk][u] * swt;
				out[9 + blk] += s[2 + blk][u] * swt;
				out[13] += 2.0 * s[2 + blk][u] * e[2 + blk][u] * swt;
			}
		}
	}
}

/* mark3mg.c:61-145 (topography, fluid markers and the non-Newtonian
 * strain-rate correction left out: they do not change eta or e of the
 * Newtonian rock types) */
void or_mk_props(long n, const float *x, const float *y, const float *z, const signed char *type, const float *tk,
		 float *eta, float *e, const or_rock_t *rocks, const float *const *st12, const double *pr,
		 const double *gx, long nx, const double *gy, long ny, const double *gz, long nz, double nubeg,
		 double nuend, double dt, double healrate)
{
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	long i;
	for (i = 0; i < n; i++) {
		int mt = type[i];
		if (!((x[i] >= 0 && y[i] >= 0 && z[i] >= 0 && x[i] <= xs && y[i] <= ys && z[i] <= zs) || mt < 50 || mt >= 100))
			continue;
		double o[14];
		or_epscalc(st12, pr, gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], o);
		double eii = pow(0.5 * (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) + o[3] * o[3] + o[4] * o[4] + o[5] * o[5], 0.5);
		double mpb = o[6] * 1e-5, mtk = tk[i];
		if (mt < 50 && mtk > 0) {
			eta[i] = (float)or_visc(&rocks[mt], mtk, mpb, nubeg, nuend);
			/* viscalc marks the marker as not yielding (its brittle
			 * inverse viscosity is zero for these rocks,
			 * mark3mg.c:1179, 1648-1659) */
			e[i] = -fabsf(e[i]);
			if (e[i] > 0) e[i] += dt * eii;
			if (e[i] > 0) {
				e[i] -= dt * healrate;
				if (e[i] < 0) e[i] = 0;
			} else {
				e[i] += dt * healrate;
				if (e[i] > 0) e[i] = 0;
			}
		}
	}
}

/* ---- full rheology, air/water, phase densities, fluid and immobile
 * markers (test infrastructure) ------------------------------------------ */

/* viscalc (mark3mg.c:1114-1681): brittle Mohr-Coulomb strength weakened by
 * the accumulated strain (:1148-1199), ductile branches A (:1266-1275),
 * B (:1277-1299), C (:1301-1326), D (:1328-1358), Karato & Wu F
 * (:1360-1378) and the grain-size model G (:1380-1623, the disabled
 * stress bisection left out), the yield flag on the strain (:1648-1659),
 * the dislocation share (:1661) and the limits (:1669-1681).  The
 * Peierls branch (:1205) is disabled in the reference and the melt
 * pathway (:1148) reads a buffer that only the disabled melt extraction
 * fills, so both are left out.  sbrit is read by G without being set when
 * the brittle branch did not run; it is taken as +inf there (no yield
 * cap).  me, mrr, mrn: the marker's marke, markrr, markrn (in/out). */
double or_viscalc(const or_rockx_t *R, int mm2, double mtk, double mpb, double eii, double sii, double rii,
		  double dis, float *me, float *mrr, float *mrn, const or_visprm_t *P)
{
	double rt = 8.314 * mtk, epsin = eii, sigin = sii;
	double strain = fabsf(*me);
	double lamb = 1.0, abrit = R->a0, bbrit = R->b0;
	if (strain > R->e1) {
		abrit = R->a1;
		bbrit = R->b1;
	} else if (strain > R->e0 && R->e1 > R->e0) {
		abrit = R->a0 + (R->a1 - R->a0) * (strain - R->e0) / (R->e1 - R->e0);
		bbrit = R->b0 + (R->b1 - R->b0) * (strain - R->e0) / (R->e1 - R->e0);
	}
	double mpf = mpb * 1e+5, nubrit = 0, sbrit = INFINITY;
	if (epsin > 0 && (bbrit > 0 || abrit * lamb > 0) && mm2 > 1) {
		sbrit = abrit + bbrit * mpf * lamb;
		if (mpf < 0) sbrit = abrit + mpf;
		if (sbrit > 0) nubrit = 1.0 / (0.5 * sbrit / epsin / rii);
		else { nubrit = 1.0 / R->n0; sbrit = 0; }
	}
	double nunewt = 0, nupowl = 0, e1, k1, sduct;
	if (R->nu0 != 0) {
		if (R->dh == 0 && R->dv == 0 && (R->ss == 0 || R->mm == 1.0)) nunewt = 1.0 / R->nu0;
		if ((R->dh != 0 || R->dv != 0) && (R->ss == 0 || R->mm == 1.0)) {
			e1 = (R->dh + R->dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			if (R->dh < 0 && R->dv > 0) { e1 = R->dh * (mtk - R->dv); if (e1 < -150.0) e1 = -150.0; }
			if (R->dh < 0 && R->dv < 0) {
				e1 = (-R->dh) * (1.0 - (mtk - (-R->dv)) / (-R->dv)) / 8.314 / (-R->dv);
				if (e1 > 150.0) e1 = 150.0;
			}
			nunewt = 1.0 / (R->nu0 * exp(e1));
		}
		if ((R->dh == 0 && R->dv == 0) && R->ss != 0 && R->mm != 1.0) {
			k1 = R->nu0 / pow(R->ss, R->mm - 1.0);
			nunewt = 1.0 / (0.5 * k1);
			sduct = pow(R->nu0 * epsin, 1.0 / R->mm);
			if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
		}
		if (R->nu0 > 0 && R->dh > 0 && (R->dh != 0 || R->dv != 0) && R->ss != 0 && R->mm != 1.0) {
			e1 = (R->dh + R->dv * mpb) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(e1);
			k1 = R->nu0 / pow(R->ss, R->mm - 1.0);
			nunewt = 1.0 / (0.5 * k1 * e1);
			sduct = pow(R->nu0 * e1 * epsin, 1.0 / R->mm);
			if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
		}
		if (R->nu0 < 0) {
			double mnu2 = 1.0;
			e1 = (R->ss + mpb * R->ll) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(e1);
			nunewt = 1.0 / (0.5 * pow(mnu2 * 1e-3 / 5e-10, 2.5) * 8e+10 / R->s1 * e1);
			nupowl = 0;
			if (epsin > 0) {
				e1 = (R->dh + R->dv * mpb) / rt / R->mm;
				if (e1 > 150.0) e1 = 150.0;
				e1 = exp(e1);
				nupowl = 1.0 / (0.5 * pow(epsin, 1.0 / R->mm - 1.0) * 8e+10 / pow(-R->nu0, 1.0 / R->mm) * e1);
			}
		}
		if (R->dh < 0) {
			double dissip, r, r0, dtsum, drlim, dt, dt0, n, AA, BB, m, pival;
			epsin = eii;
			sigin = sii;
			dissip = dis * (double)*mrn;
			pival = 2.0 * asin(1.0);
			r = r0 = *mrr;
			if (r0 <= 0) {
				*mrr = (float)(1e-3 * 2.0 / pival);
				r = r0 = *mrr;
			}
			dtsum = 0;
			drlim = 0.1;
			dt0 = P->timestep;
			n = R->mm;
			e1 = (-R->dh + mpb * R->dv) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(-e1);
			AA = R->nu0 * e1;
			m = 3.0;
			e1 = (R->ss + mpb * R->ll) / rt;
			if (e1 > 150.0) e1 = 150.0;
			e1 = exp(-e1);
			BB = R->s1 * e1;
			BB = BB / pow(pival / 2.0, m);
			if (P->timesum > P->convend) {
				int j = 0;
				do {
					double q = 4.0, p = 2.0, phi1 = 0.4, phi2 = 1.0 - phi1, eta = 3.0 * phi1 * phi2, gammaI = 1.0;
					double Eggrow = 300e+3, Aggrow = 2e+4 * pow(1e-6, p), Vggrow = R->ll, GG, Gfac, GI, f0, fG,
					       fI, Psi, DlnrDt;
					e1 = (Eggrow + mpb * Vggrow) / rt;
					if (e1 > 150.0) e1 = 150.0;
					e1 = exp(-e1);
					GG = Aggrow * e1;
					Gfac = 100.0;
					GI = GG / Gfac * q / p * pow(r0, q - p);
					f0 = 1e-3;
					e1 = 2.0 * pow(mtk / 1000, 2.9) - 2.0;
					if (e1 > 150.0) e1 = 150.0;
					e1 = exp(-e1);
					fG = f0 * e1;
					fI = fG;
					Psi = dissip;
					(void)phi2;
					DlnrDt = (eta * GI / q / pow(r0, q - 1.0) - fI * r0 * r0 / gammaI / eta * Psi) / r0;
					dt = dt0 - dtsum;
					if (fabs(DlnrDt * dt) > drlim) dt = dt * drlim / fabs(DlnrDt * dt);
					r = r0 * exp(DlnrDt * dt);
					if (r > 1e-2) r = 1e-2;
					if (r < 1e-6) r = 1e-6;
					j = j + 1;
					dtsum = dtsum + dt;
					r0 = r;
				} while (dtsum < dt0 && j < 50);
				*mrr = (float)r;
			}
			nunewt = pow(r, m) / BB / 2.0;
			nupowl = 0;
			if (epsin > 0) {
				double epsinpowl = epsin, siginpowl = pow(epsinpowl / AA, 1.0 / n);
				nupowl = siginpowl / 2.0 / epsinpowl;
				if (sigin > sbrit) nupowl = 2.0 * AA * pow(sbrit, n - 1);
				else nupowl = 1.0 / nupowl;
			}
			nunewt = 1.0 / nunewt;
		}
	}
	double nueff = nunewt + nupowl;
	if (0 > nueff) nueff = 0;
	if (nubrit > nueff) {
		nueff = nubrit;
		if (*me == 0) *me = 1e-20f;
		*me = fabsf(*me);
	} else {
		*me = -fabsf(*me);
	}
	if (R->dh < 0 && P->timesum > 1e+11) *mrn = (float)(nupowl / nueff);
	if (nueff <= 0) nueff = 1.0 / R->n1;
	nueff = 1.0 / nueff;
	if (nueff < R->n0) nueff = R->n0;
	if (nueff > R->n1) nueff = R->n1;
	if (nueff < P->nubeg) nueff = P->nubeg;
	if (nueff > P->nuend) nueff = P->nuend;
	return nueff;
}

static double inv2(const double *o, int b)
{
	return pow(0.5 * (o[b] * o[b] + o[b + 1] * o[b + 1] + o[b + 2] * o[b + 2]) + o[b + 3] * o[b + 3] +
			   o[b + 4] * o[b + 4] + o[b + 5] * o[b + 5],
		   0.5);
}

/* movemark's property update (mark3mg.c:61-145) with the full viscalc:
 * for the markers that move (inside the grid, or outgrid != 1 and type <
 * 50 or >= 100), epscalc1 at the marker; for type < 50 with T > 0 the
 * strain-rate correction of numerical diffusion (:97-105, markers away
 * from the three outer grid lines with a viscosity), viscalc, the strain
 * growth when yielding and the healing.  The topography of the marker
 * (markto, :75-88) is left to the caller. */
void or_mk_props_x(long n, const float *x, const float *y, const float *z, const signed char *type, const float *tk,
		   float *eta, float *e, float *rr, float *rn, const or_rockx_t *rocks, const float *const *st12,
		   const double *pr, const double *gx, long nx, const double *gy, long ny, const double *gz, long nz,
		   const or_visprm_t *P, double healrate, int outgrid)
{
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	long i;
	for (i = 0; i < n; i++) {
		int mt = type[i];
		if (!((x[i] >= 0 && y[i] >= 0 && z[i] >= 0 && x[i] <= xs && y[i] <= ys && z[i] <= zs) ||
		      (outgrid != 1 && (mt < 50 || mt >= 100))))
			continue;
		double o[14];
		or_epscalc(st12, pr, gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], o);
		double eii = inv2(o, 0), sii = inv2(o, 7), dis = o[13];
		double mpb = o[6] * 1e-5, mtk = tk[i];
		if (mt < 50 && mtk > 0) {
			double rii = 1.0;
			if (eta[i] > 0 && x[i] > gx[3] && x[i] < gx[nx - 4] && y[i] > gy[3] && y[i] < gy[ny - 4] && z[i] > gz[3] &&
			    z[i] < gz[nz - 4])
				rii *= pow(0.5 * sii / eii / eta[i], 0.87);
			double mnu = or_viscalc(&rocks[mt], mt, mtk, mpb, eii, sii, rii, dis, &e[i], &rr[i], &rn[i], P);
			eta[i] = (float)mnu;
			if (e[i] > 0) e[i] += P->timestep * eii;
			if (e[i] > 0) {
				e[i] -= P->timestep * healrate;
				if (e[i] < 0) e[i] = 0;
			} else {
				e[i] += P->timestep * healrate;
				if (e[i] > 0) e[i] = 0;
			}
		}
	}
}

/* dencalc1 (mark3mg.c:3006-3024); *bb = eps1[20] */
static double or_dencalc(const or_rockx_t *r, int densimod, double mtk, double mpb, double *bb)
{
	*bb = 0;
	if (densimod == 0) return r->ro;
	*bb = r->bro / (1.0 - r->bro * (mtk - 298.15));
	return r->ro * (1.0 - r->bro * (mtk - 298.15)) * (1.0 + r->aro * (mpb - 1.0) * 1e-3);
}

/* density factor of the phase transitions of ronurecalc (mark3mg.c:
 * 530-641): basalt -> garnet granulite -> eclogite, coesite -> stishovite
 * and perovskite for types 7-8 and 16-18; olivine -> spinel and
 * perovskite for 9-14; both damped below teclmax and off below teclmin */
static double or_phase_rokf(int mm2, double mtk, double mpb, double teclmin, double teclmax)
{
	double rokf = 1.0;
	int ecl = (mm2 >= 7 && mm2 <= 8) || (mm2 >= 16 && mm2 <= 18), man = mm2 >= 9 && mm2 <= 14;
	if (!ecl && !man) return 1.0;
	if (ecl) {
		double p_ga_in = -9222.0 + mtk * 14.0, p_pl_out, p_st_in, p_sp_out, p_pv_in;
		if (p_ga_in < 2000.0) p_ga_in = 2000.0;
		p_pl_out = -1460.0 + mtk * 20.0;
		if (p_pl_out < 12000.0) p_pl_out = 12000.0;
		if (mpb > p_ga_in) {
			if (mpb >= p_pl_out) rokf *= 1.16;
			else rokf *= 1.0 + 0.16 * (mpb - p_ga_in) / (p_pl_out - p_ga_in);
		}
		p_st_in = 59100.0 + mtk * 22.6;
		if (mpb > p_st_in) rokf *= 1.06;
		p_sp_out = 354000.0 - mtk * 40.0;
		p_pv_in = 352000.0 - mtk * 40.0;
		if (mpb > p_pv_in) {
			if (mpb >= p_sp_out) rokf *= 1.08;
			else rokf *= 1.0 + 0.08 * (mpb - p_pv_in) / (p_sp_out - p_pv_in);
		}
	} else {
		double p_ol_out = 91000.0 + mtk * 27.0, p_sp_in = 66000.0 + mtk * 39.0, p_sp_out, p_pv_in;
		if (p_sp_in > p_ol_out - 2000.0) p_sp_in = p_ol_out - 2000.0;
		if (mpb > p_sp_in) {
			if (mpb >= p_ol_out) rokf *= 1.06;
			else rokf *= 1.0 + 0.06 * (mpb - p_sp_in) / (p_ol_out - p_sp_in);
		}
		p_sp_out = 304000.0 - mtk * 40.0;
		p_pv_in = 302000.0 - mtk * 40.0;
		if (mpb > p_pv_in) {
			if (mpb >= p_sp_out) rokf *= 1.11;
			else rokf *= 1.0 + 0.11 * (mpb - p_pv_in) / (p_sp_out - p_pv_in);
		}
	}
	if (mtk < teclmax) {
		if (mtk < teclmin) rokf = 1.00;
		else rokf = 1.0 + (rokf - 1.0) * (mtk - teclmin) / (teclmax - teclmin);
	}
	return rokf;
}

/* water dissolved in the melt and the solidus depression it gives (Katz
 * et al. 2003 as mark3mg.c:1950-1990 uses it): xh2o = mwa / (0.01 + f
 * (1 - 0.01)) capped by the saturation 12 P^0.6 + P (P in GPa);
 * dT = 43 xh2o^0.75 */
static double or_wet_dt(double mwa, double f, double mpg)
{
	double sat = 12.0 * pow(mpg, 0.6) + 1.0 * mpg;
	double xh2o = mwa / (0.01 + f * (1.0 - 0.01));
	if (xh2o > sat) xh2o = sat;
	return 43.0 * pow(xh2o, 0.75);
}

/* meltpart11 (mark3mg.c:1872-2070).  Rock families by type: sediments and
 * upper crust (wet granite solidus, dry granite liquidus, 300 kJ/kg),
 * basalt (wet solidus of Schmidt & Poli, tholeiite liquidus, 380 kJ/kg),
 * gabbro (dry tholeiite solidus and liquidus, 380 kJ/kg), peridotite
 * (Katz et al. 2003 with cpx exhaustion; the water solidus depression by
 * bisection on dT over the melt fraction, at most 100 steps; 400 kJ/kg;
 * dry peridotite does not melt above maxpmelt).  The melt fraction is then
 * linear between the family's solidus and (effective) liquidus. */
void or_meltpart11(int mm2, double mtk, double mpb, double mwa, double maxpmelt, double *xmelt, double *hlat)
{
	double ykm = mpb * 3e-3, ts = 0, tl = 0, tll = 0, hl = 0;
	*xmelt = 0;
	*hlat = 0;
	if (ykm > 0) {
		int fam = -1;
		switch (mm2) {
		case 3: case 4: case 5: case 15: case 17: case 23: case 24: case 25: case 35: case 37: fam = 0; break;
		case 7: case 16: case 18: case 27: case 36: case 38: fam = 1; break;
		case 6: case 8: case 26: case 28: fam = 2; break;
		case 9: case 10: case 11: case 12: case 14: case 29: case 30: case 31: case 32: case 34: fam = 3; break;
		default: break;
		}
		if (fam == 0) {
			ts = ykm < 36.0 ? 889.0 + 536.6 / (ykm + 1.609) + 18.21 / (ykm + 1.609) / (ykm + 1.609)
					: 831.3 + 2.0 * ykm;
			tl = 1262.0 + 3.0 * ykm;
			hl = 300000.0;
		} else if (fam == 1) {
			ts = ykm < 48.0 ? 972.6 - 2111.0 / (ykm + 10.63) + 70033.0 / (ykm + 10.63) / (ykm + 10.63)
					: 935.4 + 0.1162 * ykm + 0.006937 * ykm * ykm;
			tl = 1423.15 + 3.5 * ykm;
			hl = 380000.0;
		} else if (fam == 2) {
			ts = 1327.15 + 3.02 * ykm;
			tl = 1423.15 + 3.5 * ykm;
			hl = 380000.0;
		} else if (fam == 3 && !(mwa <= 0 && mpb * 1e+5 > maxpmelt)) {
			double mpg = mpb * 1e-4, lo = 0, hi = 0, dt;
			int steps = 1, it;
			if (mwa > 0) {
				dt = or_wet_dt(mwa, 0.0, mpg);
				ts = 273.15 + 1085.7 + 132.9 * mpg - 5.1 * mpg * mpg - dt;
				lo = hi = dt;
				if (mtk >= ts) {
					lo = or_wet_dt(mwa, 1.0, mpg);
					steps = 100;
				}
			}
			for (it = 0; it < steps; it++) {
				double f = 0;
				dt = (lo + hi) / 2.0;
				ts = 273.15 + 1085.7 + 132.9 * mpg - 5.1 * mpg * mpg - dt;
				tll = 273.15 + 1475.0 + 80.0 * mpg - 3.2 * mpg * mpg - dt;
				tl = 273.15 + 1780.0 + 45.0 * mpg - 2.0 * mpg * mpg - dt;
				if (mtk >= ts) {
					if (mtk >= tl) {
						f = 1.0;
					} else {
						double fcpx = 0.15 / (0.5 + 0.08 * mpg);
						double tcpx = pow(fcpx, 1.0 / 1.5) * (tll - ts) + ts;
						f = mtk > tcpx ? fcpx + (1.0 - fcpx) * pow((mtk - tcpx) / (tl - tcpx), 1.5)
							       : pow((mtk - ts) / (tll - ts), 1.5);
						tl = (mtk - ts) / f + ts;
					}
				}
				if (mwa > 0) {
					if (or_wet_dt(mwa, f, mpg) > dt) lo = dt;
					else hi = dt;
				}
				if (fabs(hi - lo) < 0.001) break;
			}
			hl = 400000.0;
		}
	}
	if (tl != 0) {
		double f = (mtk - ts) / (tl - ts);
		if (f < 0) f = 0;
		if (f > 1.0) f = 1.0;
		*xmelt = f;
		*hlat = hl * f;
	}
}

/* meltpart0 (mark3mg.c:2152-2340).  e[0..6] = eps1[20..26]: thermal
 * expansion term, melt fraction, latent heat, density, (unused), heat
 * capacity, conductivity of the partially molten rock: melt (type mm2) and
 * solid (mm2 - 20, type 11 for 34) mixed by the melt fraction less the
 * extracted melt; densimod 1 adds the solid's eclogite (27-28, 36-38) and
 * olivine-spinel-perovskite (29-34) factors; inside the melting interval
 * the latent heat's pressure and temperature derivatives (central
 * differences over 300 bar and 10 K) are added to the expansion term
 * (capped at 1e-2) and the heat capacity (100..5e4).  e[1] and e[2] are
 * those of the last melt-fraction evaluation, as in eps1. */
int or_meltpart0(const or_rockx_t *rocks, int mm2, double mtk, double mpb, double mwa, double mex, int densimod,
		 double teclmin, double teclmax, double maxpmelt, double *e)
{
	const or_rockx_t *A, *B;
	double f, xm, hl, rokf = 1.0, a, b;
	int q, solid;
	for (q = 0; q < 7; q++) e[q] = 0;
	if (mm2 <= 21) return 0;
	or_meltpart11(mm2, mtk, mpb, mwa, maxpmelt, &xm, &hl);
	e[1] = xm;
	e[2] = hl;
	f = xm - mex;
	if (f < 0) f = 0;
	solid = mm2 == 34 ? 11 : mm2 - 20;
	A = &rocks[mm2];
	B = &rocks[solid];
	e[0] = (A->bro * f + B->bro * (1.0 - f)) / (1.0 - (A->bro * f + B->bro * (1.0 - f)) * (mtk - 298.15));
	if (densimod == 0) {
		e[3] = A->ro * f + B->ro * (1.0 - f);
	} else {
		if ((mm2 >= 27 && mm2 <= 28) || (mm2 >= 36 && mm2 <= 38)) rokf = or_phase_rokf(7, mtk, mpb, teclmin, teclmax);
		if (mm2 >= 29 && mm2 <= 34) rokf = or_phase_rokf(9, mtk, mpb, teclmin, teclmax);
		e[3] = f * A->ro * (1.0 - A->bro * (mtk - 298.15)) * (1.0 + A->aro * (mpb - 1.0) * 1e-3) +
		       (1.0 - f) * rokf * B->ro * (1.0 - B->bro * (mtk - 298.15)) * (1.0 + B->aro * (mpb - 1.0) * 1e-3);
	}
	e[5] = A->cp * f + B->cp * (1.0 - f);
	e[6] = ((A->kt + A->kf / (mtk + 77.0)) * exp(A->kp * mpb)) * f +
	       ((B->kt + B->kf / (mtk + 77.0)) * exp(B->kp * mpb)) * (1.0 - f);
	if (f > 0 && f < 1.0) {
		double dp = 300.0, dT = 10.0, d;
		or_meltpart11(mm2, mtk, mpb - dp, mwa, maxpmelt, &xm, &a);
		or_meltpart11(mm2, mtk, mpb + dp, mwa, maxpmelt, &xm, &b);
		d = (a - b) * (e[3] / (mtk * 2.0 * dp * 1e+5));
		e[0] += d;
		if (e[0] < -1e-2) e[0] = -1e-2;
		else if (e[0] > 1e-2) e[0] = 1e-2;
		or_meltpart11(mm2, mtk + dT, mpb, mwa, maxpmelt, &xm, &a);
		or_meltpart11(mm2, mtk - dT, mpb, mwa, maxpmelt, &xm, &b);
		e[1] = xm;
		e[2] = b;
		e[5] += (a - b) / (2.0 * dT);
		if (e[5] < 1e+2) e[5] = 1e+2;
		else if (e[5] > 5e+4) e[5] = 5e+4;
	}
	return 1;
}

/* ronurecalc (mark3mg.c:404-1015) with the full rheology: erosion,
 * sedimentation and water/air retyping (:465-472), viscalc from epscalc1
 * strain rates when the marker has no viscosity or timestep <= 0
 * (:506-520), dencalc1 with the phase-transition density factors
 * (:521-641), the water/air proportions of markers at the water level
 * (:723-763; the y-cell of the marker sets the marker spacing, where the
 * reference reads a shared index array), the viscosity caps, then the
 * scatter of or_ronurecalc.  Melt types (> 20) take the partially molten
 * rock's properties (or_meltpart0) from their water content mw and
 * extracted melt mex (NULL: zero).  densimod >= 2 is not restated: returns
 * -1 if a participating marker needs it.  st12 may be NULL (zero strain
 * rates and stresses). */
int or_ronurecalc_x(long n, const float *x, const float *y, const float *z, const float *tk, float *eta,
		    signed char *type, float *e, float *rr, float *rn, const double *gx, long nx, const double *gy,
		    long ny, const double *gz, long nz, const double *pr, const float *const *st12,
		    const or_rockx_t *rocks, const or_visprm_t *P, const or_mkx_t *M, const float *mw,
		    const float *mex, float *fld)
{
	long N = nx * ny * nz, i, q;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	float *acc = calloc((size_t)(11 * N), sizeof(float));
	double *accd = calloc((size_t)(11 * N), sizeof(double));
	double *w = calloc((size_t)(6 * N), sizeof(double));
	int rc = 0;
	for (i = 0; i < n && !rc; i++) {
		long cell[3], m1, m2, m3, m10, m20, m30;
		double mpb, mtk, mnu, mro, mbb, mcp, mkt, mht, dx, dy, dz, eii = 0, sii = 0, dis = 0;
		int mm2 = type[i];
		if (!(x[i] > 0 && y[i] > 0 && z[i] > 0 && x[i] < xs && y[i] < ys && z[i] < zs && mm2 < 50 && tk[i] > 0))
			continue;
		if (mm2 > 1 && y[i] < M->eroslev) mm2 = type[i] = 0;
		if (mm2 < 2 && y[i] > M->sedilev) mm2 = type[i] = (signed char)M->sedtype;
		if (mm2 < 2 && y[i] > M->waterlev) mm2 = type[i] = 1;
		if (mm2 < 2 && y[i] < M->waterlev) mm2 = type[i] = 0;
		if (M->densimod >= 2) { rc = -1; break; }
		mpb = or_prcalc(pr, gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], cell) * 1e-5;
		mtk = tk[i];
		mnu = eta[i];
		if (mnu <= 0 || P->timestep <= 0) {
			double o[14] = {0};
			if (st12) or_epscalc(st12, pr, gx, nx, gy, ny, gz, nz, x[i], y[i], z[i], o);
			eii = inv2(o, 0); sii = inv2(o, 7); dis = o[13];
			mnu = or_viscalc(&rocks[mm2], mm2, mtk, mpb, eii, sii, 1.0, dis, &e[i], &rr[i], &rn[i], P);
			eta[i] = (float)mnu;
		}
		mro = or_dencalc(&rocks[mm2], M->densimod, mtk, mpb, &mbb);
		mro *= or_phase_rokf(mm2, mtk, mpb, M->teclmin, M->teclmax);
		mcp = rocks[mm2].cp;
		mkt = (rocks[mm2].kt + rocks[mm2].kf / (mtk + 77.0)) * exp(rocks[mm2].kp * mpb);
		mht = rocks[mm2].ht;
		if (mm2 > 20) {
			/* partially molten rock (mark3mg.c:660-686) */
			double e7[7];
			or_meltpart0(rocks, mm2, mtk, mpb, mw ? (double)mw[i] : 0.0, mex ? (double)mex[i] : 0.0,
				     M->densimod, M->teclmin, M->teclmax, M->maxpmelt, e7);
			mro = e7[3];
			mbb = e7[0];
			mcp = e7[5];
			mkt = e7[6];
			if ((mm2 == 27 || mm2 == 28) && e7[1] > 0.001) {
				/* basaltic melt (Schott and Schmeling 1995) */
				double nm = 1e+15 * exp(2.5 + pow((1.0 - e7[1]) / e7[1], 0.48) * (1.0 - e7[1]));
				if (nm < mnu) {
					mnu = nm;
					if (mnu < rocks[mm2].n0) mnu = rocks[mm2].n0;
					if (mnu > rocks[mm2].n1) mnu = rocks[mm2].n1;
					if (mnu < P->nubeg) mnu = P->nubeg;
					if (mnu > P->nuend) mnu = P->nuend;
					eta[i] = (float)mnu;
				}
			}
		}
		if (M->waterlev >= 0 && mm2 < 2) {
			double dywa = (gy[cell[1] + 1] - gy[cell[1]]) / 2.0 / ((double)M->mnumy);
			double lev = ((double)y[i] - (M->waterlev - dywa)) / (2.0 * dywa);
			if (lev < 1.0 && lev > 0) {
				double abb, wbb, anu, aro, wnu, wro;
				anu = or_viscalc(&rocks[0], 0, mtk, mpb, eii, sii, 1.0, dis, &e[i], &rr[i], &rn[i], P);
				aro = or_dencalc(&rocks[0], M->densimod, mtk, mpb, &abb);
				wnu = or_viscalc(&rocks[1], 1, mtk, mpb, eii, sii, 1.0, dis, &e[i], &rr[i], &rn[i], P);
				wro = or_dencalc(&rocks[1], M->densimod, mtk, mpb, &wbb);
				mnu = wnu * lev + anu * (1.0 - lev);
				mro = wro * lev + aro * (1.0 - lev);
				mbb = wbb * lev + abb * (1.0 - lev);
				mcp = rocks[1].cp * lev + rocks[0].cp * (1.0 - lev);
				mkt = rocks[1].kt * lev + rocks[0].kt * (1.0 - lev);
				mht = rocks[1].ht * lev + rocks[0].ht * (1.0 - lev);
			}
		}
		if (y[i] > M->vdeep && mnu > M->nudeep) mnu = M->nudeep;
		if (y[i] < 5e+4 && x[i] > xs - 1e+4 && z[i] > 1.5e+5 && z[i] < 8.7e+5 && mnu > 1e+19) mnu = 1e+19;
		m10 = cell[0]; m20 = cell[1]; m30 = cell[2];
		dx = gx[m10 + 1] - gx[m10]; dy = gy[m20 + 1] - gy[m20]; dz = gz[m30 + 1] - gz[m30];
		for (m3 = m30; m3 <= m30 + 1; m3++) {
			double ddz = 1.0 - fabs((gz[m3] - z[i]) / dz);
			for (m1 = m10; m1 <= m10 + 1; m1++) {
				double ddx = 1.0 - fabs((gx[m1] - x[i]) / dx);
				for (m2 = m20; m2 <= m20 + 1; m2++) {
					double ddy = 1.0 - fabs((gy[m2] - y[i]) / dy);
					double swt = ddx * ddy * ddz, s1;
					long m4 = m3 * nx * ny + m1 * ny + m2;
					if (ddx > M->nukoef && ddy > M->nukoef && ddz > M->nukoef) {
						or_acc_add(acc, accd, m4, mnu * swt);
						w[m4] += swt;
					}
					if (m1 > m10 && m2 > m20 && m3 > m30) {
						s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (0.5 - fabs(ddy - 0.5)) * (0.5 - fabs(ddz - 0.5));
						or_acc_add(acc, accd, 7 * N + m4, mnu * s1);
						w[2 * N + m4] += s1;
					}
					if (m3 == m30 && ddx > 0.5 && ddy > 0.5) {
						s1 = 8.0 * (ddx - 0.5) * (ddy - 0.5) * (0.5 - fabs(ddz - 0.5));
						or_acc_add(acc, accd, 8 * N + m4, mnu * s1);
						w[3 * N + m4] += s1;
					}
					if (m2 == m20 && ddx > 0.5 && ddz > 0.5) {
						s1 = 8.0 * (ddx - 0.5) * (0.5 - fabs(ddy - 0.5)) * (ddz - 0.5);
						or_acc_add(acc, accd, 9 * N + m4, mnu * s1);
						w[4 * N + m4] += s1;
					}
					if (m1 == m10 && ddy > 0.5 && ddz > 0.5) {
						s1 = 8.0 * (0.5 - fabs(ddx - 0.5)) * (ddy - 0.5) * (ddz - 0.5);
						or_acc_add(acc, accd, 10 * N + m4, mnu * s1);
						w[5 * N + m4] += s1;
					}
					if (swt > 0) {
						or_acc_add(acc, accd, 1 * N + m4, mro * swt);
						or_acc_add(acc, accd, 2 * N + m4, mcp * mro * swt);
						or_acc_add(acc, accd, 3 * N + m4, mkt * swt);
						or_acc_add(acc, accd, 4 * N + m4, mht * swt);
						or_acc_add(acc, accd, 5 * N + m4, mbb * swt);
						or_acc_add(acc, accd, 6 * N + m4, mtk * swt);
						w[N + m4] += swt;
					}
				}
			}
		}
	}
	if (!rc)
		for (i = 0; i < N; i++) {
			if (w[i] > 0.01) fld[i] = (float)(or_acc_get(acc, accd, i) / w[i]);
			for (q = 0; q < 4; q++)
				if (w[(2 + q) * N + i] > 0.01) fld[(7 + q) * N + i] = (float)(or_acc_get(acc, accd, (7 + q) * N + i) / w[(2 + q) * N + i]);
			if (w[N + i] > 0.01) {
				fld[N + i] = (float)(or_acc_get(acc, accd, N + i) / w[N + i]);
				fld[2 * N + i] = (float)(or_acc_get(acc, accd, 2 * N + i) / w[N + i] / fld[N + i]);
				for (q = 3; q <= 6; q++) fld[q * N + i] = (float)(or_acc_get(acc, accd, q * N + i) / w[N + i]);
			}
			if (fld[i] < P->nubeg) fld[i] = (float)P->nubeg;
			if (fld[i] > P->nuend) fld[i] = (float)P->nuend;
		}
	free(acc);
	free(accd);
	free(w);
	return rc;
}

/* movemark's motion (mark3mg.c:61, 146-284): the markers that move (inside
 * the grid, or outgrid != 1 and type < 50 or >= 100) take the velocity of
 * vxyzcalc1 (markmod 1) or of the four RK4 stages, plus for fluid markers
 * (types 50-99) the water velocity: vymelt above density 1100 (markd),
 * else vyfluid; with vyfluid > 0 and T 0 or above 298 K it is turned along
 * the pressure gradient of the eight P nodes around the marker (:157-201,
 * the y and z offsets scaled by the x factor and the upper nodes 6, 7
 * shifted by one y line as in the reference; a node index past the last
 * node reads 0), else it is -|vy_water| (:203-207).  outgrid 2 keeps only
 * the motion normal to a face the marker is outside of (:245-251).
 * Markers below 100 take the depth rules and the out-of-grid reset;
 * immobile markers (>= 100) only move. */
void or_rk4_x(long n, float *x, float *y, float *z, float *tk, float *eta, signed char *type, const float *md,
	      const float *vx, const float *vy, const float *vz, const double *pr, const double *gx, long nx,
	      const double *gy, long ny, const double *gz, long nz, double dt, double zdeep, double vdeep,
	      double tdeep, const or_movex_t *M)
{
	long i, N = nx * ny * nz;
	int q;
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	for (i = 0; i < n; i++) {
		double v1[3], v2[3], v3[3], v4[3], v[3], px = x[i], py = y[i], pz = z[i], vw[3] = {0, 0, 0};
		int mt = type[i];
		if (!((x[i] >= 0 && y[i] >= 0 && z[i] >= 0 && x[i] <= xs && y[i] <= ys && z[i] <= zs) ||
		      (M->outgrid != 1 && (mt < 50 || mt >= 100))))
			continue;
		if (mt >= 50 && mt < 100) {
			vw[1] = M->vyfluid;
			if (md && md[i] > 1100.0) vw[1] = M->vymelt;
			if (M->vyfluid > 0 && (tk[i] == 0 || tk[i] > 298.0)) {
				long c0 = serch(gx, nx, clampx(px, xs)), c1 = serch(gy, ny, clampx(py, ys)),
				     c2 = serch(gz, nz, clampx(pz, zs));
				long m1 = c0, m2 = c1, m3 = c2, p[8];
				double xkf, ykf, zkf, ax, ay, az, d[3], k[3], pv[8];
				if (x[i] > (gx[m1] + gx[m1 + 1]) / 2.0) m1 += 1;
				if (m1 < 1) m1 = 1;
				if (m1 > nx - 2) m1 = nx - 2;
				if (y[i] > (gy[m2] + gy[m2 + 1]) / 2.0) m2 += 1;
				if (m2 < 1) m2 = 1;
				if (m2 > ny - 2) m2 = ny - 2;
				if (z[i] > (gz[m3] + gz[m3 + 1]) / 2.0) m3 += 1;
				if (m3 < 1) m3 = 1;
				if (m3 > nz - 2) m3 = nz - 2;
				xkf = 2.0 / (gx[m1 + 1] - gx[m1 - 1]);
				ykf = 2.0 / (gy[m2 + 1] - gy[m2 - 1]);
				zkf = 2.0 / (gz[m3 + 1] - gz[m3 - 1]);
				ax = xkf * (x[i] - (gx[m1 - 1] + gx[m1]) / 2.0);
				ay = xkf * (y[i] - (gy[m2 - 1] + gy[m2]) / 2.0);
				az = xkf * (z[i] - (gz[m3 - 1] + gz[m3]) / 2.0);
				p[0] = m3 * nx * ny + m1 * ny + m2;
				p[1] = p[0] + 1;
				p[2] = p[0] + ny;
				p[3] = p[2] + 1;
				p[4] = p[0] + nx * ny;
				p[5] = p[4] + 1;
				p[6] = p[5] + ny;
				p[7] = p[6] + 1;
				for (q = 0; q < 8; q++) pv[q] = p[q] < N ? pr[p[q]] : 0.0;
				d[0] = xkf * ((1.0 - ay) * (1.0 - az) * (pv[2] - pv[0]) + ay * (1.0 - az) * (pv[3] - pv[1]) +
					      (1.0 - ay) * az * (pv[6] - pv[4]) + ay * az * (pv[7] - pv[5]));
				d[1] = ykf * ((1.0 - ax) * (1.0 - az) * (pv[1] - pv[0]) + ax * (1.0 - az) * (pv[3] - pv[2]) +
					      (1.0 - ax) * az * (pv[5] - pv[4]) + ax * az * (pv[7] - pv[6]));
				d[2] = zkf * ((1.0 - ax) * (1.0 - ay) * (pv[4] - pv[0]) + ax * (1.0 - ay) * (pv[6] - pv[2]) +
					      (1.0 - ax) * ay * (pv[5] - pv[1]) + ax * ay * (pv[7] - pv[3]));
				for (q = 0; q < 3; q++) {
					k[q] = (1000.0 * M->g[q] - d[q]) / (2300.0 * 9.81);
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
		or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px, py, pz, v1);
		if (M->markmod == 1) {
			for (q = 0; q < 3; q++) v[q] = v1[q] + vw[q];
		} else {
			for (q = 0; q < 3; q++) v1[q] += vw[q];
			or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v1[0] * dt / 2.0, py + v1[1] * dt / 2.0, pz + v1[2] * dt / 2.0, v2);
			for (q = 0; q < 3; q++) v2[q] += vw[q];
			or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v2[0] * dt / 2.0, py + v2[1] * dt / 2.0, pz + v2[2] * dt / 2.0, v3);
			for (q = 0; q < 3; q++) v3[q] += vw[q];
			or_vxyz(vx, vy, vz, gx, nx, gy, ny, gz, nz, px + v3[0] * dt, py + v3[1] * dt, pz + v3[2] * dt, v4);
			for (q = 0; q < 3; q++) v4[q] += vw[q];
			for (q = 0; q < 3; q++) v[q] = (v1[q] + 2.0 * v2[q] + 2.0 * v3[q] + v4[q]) / 6.0;
		}
		if (M->outgrid == 2) {
			if (x[i] < 0 || (double)x[i] > xs) v[1] = v[2] = 0;
			if (y[i] < 0 || (double)y[i] > ys) v[0] = v[2] = 0;
			if (z[i] < 0 || (double)z[i] > zs) v[0] = v[1] = 0;
		}
		if (mt < 100) {
			double yold = y[i];
			if (y[i] > zdeep && v[1] < 0 && tk[i] < tdeep && type[i] == 10) tk[i] = (float)tdeep;
			if (y[i] > zdeep && v[1] < 0 && type[i] != 10) type[i] = 10;
			if (y[i] > zdeep && type[i] != 10 && type[i] != 9) type[i] = 9;
			x[i] += dt * v[0] - 0.0;
			y[i] += dt * v[1];
			z[i] += dt * v[2];
			if ((yold > vdeep && y[i] < vdeep) || (yold > zdeep && y[i] < zdeep)) {
				if (tk[i] < tdeep) tk[i] = (float)tdeep;
				type[i] = 10;
			}
			if (x[i] < 0 || y[i] < 0 || z[i] < 0 || x[i] > xs || y[i] > ys || z[i] > zs) {
				tk[i] = 0;
				eta[i] = 0;
			}
		} else {
			x[i] += dt * v[0] - 0.0;
			y[i] += dt * v[1];
			z[i] += dt * v[2];
		}
	}
}
