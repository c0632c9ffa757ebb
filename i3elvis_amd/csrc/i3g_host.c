/*
 * Host-side setup for the i3g hot path: multigrid hierarchy, boundary-rule
 * expansion and the seeded synthetic configurations.  Plain C, no device.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/i3g.h"

/* Hierarchy as the reference's gridcheck() (load3mg.c:1207-1233):
 * mgx[n] = ((mgx[n-1]-5)/2)+5, level offsets mgp accumulate mgn. */
int i3g_hierarchy(long nx, long ny, long nz, int multinum, i3g_hier *h)
{
	int l;
	if (!h || nx < 5 || ny < 5 || nz < 5 || multinum < 0 ||
	    multinum >= I3G_MAXLEV)
		return I3G_EINVAL;
	memset(h, 0, sizeof(*h));
	h->nlev = multinum + 1;
	h->mgx[0] = nx; h->mgy[0] = ny; h->mgz[0] = nz;
	h->mgp[0] = 0; h->mgs[0] = 1;
	h->mgn[0] = nx * ny * nz;
	for (l = 1; l <= multinum; l++) {
		h->mgp[l] = h->mgp[l - 1] + h->mgn[l - 1];
		h->mgs[l] = h->mgs[l - 1] * 2;
		h->mgx[l] = ((h->mgx[l - 1] - 5) / 2) + 5;
		h->mgy[l] = ((h->mgy[l - 1] - 5) / 2) + 5;
		h->mgz[l] = ((h->mgz[l - 1] - 5) / 2) + 5;
		h->mgn[l] = h->mgx[l] * h->mgy[l] * h->mgz[l];
		if (h->mgx[l] < 6 || h->mgy[l] < 6 || h->mgz[l] < 6)
			return I3G_EINVAL;
	}
	h->ntot = h->mgp[multinum] + h->mgn[multinum];
	return I3G_OK;
}

/* Coarse lines (load3mg.c:1235-1282): three outer lines per side copied,
 * interior line n2 takes fine line (n2-2)*mgs+2. */
static void level_axis(long n0, long nl, long s, const double *g, double *out)
{
	long i;
	out[0] = g[0]; out[1] = g[1]; out[2] = g[2];
	out[nl - 1] = g[n0 - 1]; out[nl - 2] = g[n0 - 2]; out[nl - 3] = g[n0 - 3];
	for (i = 3; i < nl - 3; i++) out[i] = g[(i - 2) * s + 2];
}

int i3g_level_lines(const i3g_hier *h, int l, const double *gx, const double *gy,
		    const double *gz, double *lx, double *ly, double *lz)
{
	if (!h || l < 0 || l >= h->nlev) return I3G_EINVAL;
	level_axis(h->mgx[0], h->mgx[l], h->mgs[l], gx, lx);
	level_axis(h->mgy[0], h->mgy[l], h->mgs[l], gy, ly);
	level_axis(h->mgz[0], h->mgz[l], h->mgs[l], gz, lz);
	return I3G_OK;
}

/* ------------------------------------------------------------------ */
/* Boundary rules                                                     */
/* ------------------------------------------------------------------ */

static void rule(i3g_bc_rule *r, int var, long a0, int ra0, long a1, int ra1,
		 long b0, int rb0, long b1, int rb1, long c0, int rc0, long c1,
		 int rc1, double cnst, double koef, int d1, int d2, int d3)
{
	r->var = var;
	r->m10 = a0; r->rel10 = ra0; r->m11 = a1; r->rel11 = ra1;
	r->m20 = b0; r->rel20 = rb0; r->m21 = b1; r->rel21 = rb1;
	r->m30 = c0; r->rel30 = rc0; r->m31 = c1; r->rel31 = rc1;
	r->cnst = cnst; r->koef = koef;
	r->dm1 = d1; r->dm2 = d2; r->dm3 = d3;
}

/* The boundary block of the reference deck init.t3c:115-168, in the same
 * order (later rules overwrite earlier ones on shared unknowns), with the
 * 0.99 bottom coefficients set to 1.0 (free slip) and fixed temperatures.
 * Indices: value (rel=0) or (n-1)+value (rel=1), as 'x'/'y'/'z' there. */
int i3g_bc_rules_freeslip(double t_top, double t_bottom, i3g_bc_rule *r, int cap)
{
	int n = 0;
	if (cap < 38) return I3G_EINVAL;
	/* Upper boundary (y = 0) */
	rule(&r[n++], 0, 1,0, 0,1,  1,0, 1,0,  1,0, 1,0,   0, 1.0, 0, 0, 1);
	rule(&r[n++], 0, 1,0, 1,0,  1,0, 1,0,  2,0, -1,1,  0, 1.0, 1, 0, 0);
	rule(&r[n++], 0, 1,0, 0,1,  1,0, 1,0,  0,1, 0,1,   0, 1.0, 0, 0, -1);
	rule(&r[n++], 0, 0,1, 0,1,  1,0, 1,0,  2,0, -1,1,  0, 1.0, -1, 0, 0);
	rule(&r[n++], 1, 1,0, -1,1, 0,0, 0,0,  0,0, -1,1,  0, 1.0, 0, 1, 0);
	rule(&r[n++], 2, 0,0, -1,1, 0,0, 0,0,  0,0, -1,1,  0, 0.0, 0, 0, 0);
	rule(&r[n++], 3, 0,0, -1,1, 0,0, 0,0,  1,0, -1,1,  0, 1.0, 0, 1, 0);
	rule(&r[n++], 4, 0,0, 0,1,  0,0, 0,0,  0,0, 0,1,   t_top, 0.0, 0, 0, 0);
	/* Lower boundary (y = ysize) */
	rule(&r[n++], 0, 1,0, 0,1,  0,1, 0,1,  1,0, 1,0,   0, 1.0, 0, 0, 1);
	rule(&r[n++], 0, 1,0, 1,0,  0,1, 0,1,  2,0, -1,1,  0, 1.0, 1, 0, 0);
	rule(&r[n++], 0, 1,0, 0,1,  0,1, 0,1,  0,1, 0,1,   0, 1.0, 0, 0, -1);
	rule(&r[n++], 0, 0,1, 0,1,  0,1, 0,1,  2,0, -1,1,  0, 1.0, -1, 0, 0);
	rule(&r[n++], 1, 1,0, -1,1, -1,1, -1,1, 0,0, -1,1, 0, 1.0, 0, -1, 0);
	rule(&r[n++], 2, 0,0, -1,1, 0,1, 0,1,  0,0, -1,1,  0, 0.0, 0, 0, 0);
	rule(&r[n++], 3, 0,0, -1,1, -1,1, -1,1, 1,0, -1,1, 0, 1.0, 0, -1, 0);
	rule(&r[n++], 4, 0,0, 0,1,  0,1, 0,1,  0,0, 0,1,   t_bottom, 0.0, 0, 0, 0);
	/* Left boundary (x = 0) */
	rule(&r[n++], 0, 1,0, 1,0,  2,0, -1,1, 1,0, 1,0,   0, 1.0, 1, 0, 1);
	rule(&r[n++], 0, 1,0, 1,0,  2,0, -1,1, 0,1, 0,1,   0, 1.0, 1, 0, -1);
	rule(&r[n++], 1, 0,0, 0,0,  0,0, -1,1, 0,0, -1,1,  0, 0.0, 0, 0, 0);
	rule(&r[n++], 2, 0,0, 0,0,  1,0, -1,1, 0,0, -1,1,  0, 1.0, 1, 0, 0);
	rule(&r[n++], 3, 0,0, 0,0,  0,0, -1,1, 1,0, -1,1,  0, 1.0, 1, 0, 0);
	rule(&r[n++], 4, 0,0, 0,0,  1,0, -1,1, 0,0, 0,1,   0, 1.0, 1, 0, 0);
	/* Right boundary (x = xsize) */
	rule(&r[n++], 0, 0,1, 0,1,  2,0, -1,1, 1,0, 1,0,   0, 1.0, -1, 0, 1);
	rule(&r[n++], 0, 0,1, 0,1,  2,0, -1,1, 0,1, 0,1,   0, 1.0, -1, 0, -1);
	rule(&r[n++], 1, 0,1, 0,1,  0,0, -1,1, 0,0, -1,1,  0, 0.0, 0, 0, 0);
	rule(&r[n++], 2, -1,1, -1,1, 1,0, -1,1, 0,0, -1,1, 0, 1.0, -1, 0, 0);
	rule(&r[n++], 3, -1,1, -1,1, 0,0, -1,1, 1,0, -1,1, 0, 1.0, -1, 0, 0);
	rule(&r[n++], 4, 0,1, 0,1,  1,0, -1,1, 0,0, 0,1,   0, 1.0, -1, 0, 0);
	/* Front boundary (z = 0) */
	rule(&r[n++], 1, 1,0, -1,1, 0,0, -1,1, 0,0, 0,0,   0, 1.0, 0, 0, 1);
	rule(&r[n++], 2, 0,0, -1,1, 1,0, -1,1, 0,0, 0,0,   0, 1.0, 0, 0, 1);
	rule(&r[n++], 3, 0,0, -1,1, 0,0, -1,1, 0,0, 0,0,   0, 0.0, 0, 0, 0);
	rule(&r[n++], 4, 1,0, -1,1, 1,0, -1,1, 0,0, 0,0,   0, 1.0, 0, 0, 1);
	/* Back boundary (z = zsize) */
	rule(&r[n++], 1, 1,0, -1,1, 0,0, -1,1, -1,1, -1,1, 0, 1.0, 0, 0, -1);
	rule(&r[n++], 2, 0,0, -1,1, 1,0, -1,1, -1,1, -1,1, 0, 1.0, 0, 0, -1);
	rule(&r[n++], 3, 0,0, -1,1, 0,0, -1,1, 0,1, 0,1,   0, 0.0, 0, 0, 0);
	rule(&r[n++], 4, 1,0, -1,1, 1,0, -1,1, 0,1, 0,1,   0, 1.0, 0, 0, -1);
	return n;
}

/* Coarse-level index of a level-0 rule bound (in3mg.c:323-340). */
static long map_bound(long m, long n0, long nl, long s)
{
	long r = (m - 2) / s + 2;
	if (m < 3) r = m;
	if (m > n0 - 4) r = nl - (n0 - m);
	return r;
}

long i3g_bc_expand(const i3g_bc_rule *rules, int nrules, const i3g_hier *h,
		   int *bondm, float *bondv, int *bondn, long cap)
{
	long nx = h->mgx[0], ny = h->mgy[0], nz = h->mgz[0];
	long nodenum = nx * ny * nz, bondnum = 0;
	int k, l;
	memset(bondm, 0, sizeof(int) * (size_t)(10 * nodenum));
	for (k = 0; k < nrules; k++) {
		const i3g_bc_rule *r = &rules[k];
		long m10 = r->m10 + (r->rel10 ? nx - 1 : 0);
		long m11 = r->m11 + (r->rel11 ? nx - 1 : 0);
		long m20 = r->m20 + (r->rel20 ? ny - 1 : 0);
		long m21 = r->m21 + (r->rel21 ? ny - 1 : 0);
		long m30 = r->m30 + (r->rel30 ? nz - 1 : 0);
		long m31 = r->m31 + (r->rel31 ? nz - 1 : 0);
		int shift = (r->dm1 || r->dm2 || r->dm3);
		for (l = 0; l < h->nlev; l++) {
			long mx = h->mgx[l], my = h->mgy[l], mz = h->mgz[l];
			long s = h->mgs[l];
			long a0 = map_bound(m10, nx, mx, s), a1 = map_bound(m11, nx, mx, s);
			long b0 = map_bound(m20, ny, my, s), b1 = map_bound(m21, ny, my, s);
			long c0 = map_bound(m30, nz, mz, s), c1 = map_bound(m31, nz, mz, s);
			long m1, m2, m3;
			for (m3 = c0; m3 <= c1; m3++)
			for (m1 = a0; m1 <= a1; m1++)
			for (m2 = b0; m2 <= b1; m2++) {
				long m4 = h->mgp[l] + m3 * mx * my + m1 * my + m2;
				long m6 = m4 + r->dm3 * mx * my + r->dm1 * my + r->dm2;
				long m5;
				if (m1 < 0 || m2 < 0 || m3 < 0 || m1 >= mx || m2 >= my || m3 >= mz)
					return I3G_EBC;
				if (++bondnum >= cap) return I3G_ECAP;
				if (r->var < 4) {
					m5 = m4 * 4 + r->var;
					bondm[m5] = (int)bondnum;
					bondv[2 * bondnum + 0] = (float)r->cnst;
					if (shift) {
						bondv[2 * bondnum + 1] = (float)r->koef;
						bondn[bondnum] = (int)(m6 * 4 + r->var + 1);
					} else {
						bondv[2 * bondnum + 1] = 0;
						bondn[bondnum] = 0;
					}
				} else {
					m5 = 8 * nodenum + m4;
					bondm[m5] = (int)bondnum;
					bondv[2 * bondnum + 0] = (float)r->cnst;
					if (shift) {
						bondv[2 * bondnum + 1] = (float)r->koef;
						bondn[bondnum] = (int)(m6 + 1);
					} else {
						bondv[2 * bondnum + 1] = 0;
						bondn[bondnum] = 0;
					}
				}
			}
		}
	}
	return bondnum;
}

/* ------------------------------------------------------------------ */
/* Seeded synthetic configurations                                    */
/* ------------------------------------------------------------------ */

/* splitmix64 keyed by (seed, stream, index) -> uniform [0,1) */
static uint64_t mix64(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

static double urand(uint64_t seed, uint64_t stream, uint64_t idx)
{
	uint64_t z = mix64(seed ^ mix64(stream * 0x100000001b3ULL ^ mix64(idx)));
	return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
	int which;
	double xs, ys, zs;
	int nblob;
	double bx[32], by[32], bz[32], bs[32], ba[32];
} cfg_t;

/* Material at a point: returns eta, rho, T. */
static void material(const cfg_t *c, double x, double y, double z, double *eta,
		     double *rho, double *t)
{
	if (c->which == 0) {
		double dx = x - 50e3, dy = y - 50e3, dz = z - 50e3;
		int in = dx * dx + dy * dy + dz * dz <= 20e3 * 20e3;
		*eta = in ? 1e23 : 1e21;
		*rho = in ? 3200.0 : 3300.0;
		*t = 273.0 + (1600.0 - 273.0) * y / c->ys;
	} else if (c->which == 1) {
		double b = 0, le;
		int k;
		for (k = 0; k < c->nblob; k++) {
			double dx = x - c->bx[k], dy = y - c->by[k], dz = z - c->bz[k];
			double r2 = dx * dx + dy * dy + dz * dz;
			b += c->ba[k] * exp(-r2 / (2.0 * c->bs[k] * c->bs[k]));
		}
		le = 21.0 + b;
		if (le < 19.0) le = 19.0;
		if (le > 23.0) le = 23.0;
		*eta = pow(10.0, le);
		*rho = 3300.0 - 50.0 * b;
		*t = 273.0 + (1600.0 - 273.0) * y / c->ys;
	} else {
		/* Subduction: trench at x=400 km; slab top surface depth
		 * d = x - 400 km (45 degrees) for x >= 400 km, 100 km thick
		 * (normal), down to 400 km depth, z in [200, 800] km.  A 10 km
		 * weak channel (1e17) lines the slab top above 150 km depth. */
		const double lith = 100e3, xt = 400e3, thick = 100e3;
		double d_perp = ((x - xt) - y) / sqrt(2.0); /* >0 below-right of top */
		int inz = z >= 200e3 && z <= 800e3;
		double age = 60e6 * 3.15576e7, kappa = 1e-6;
		*eta = 1e20; *rho = 3250.0;
		if (y <= lith) { *eta = 1e23; *rho = 3300.0; }
		if (inz && y > 0 && y <= 400e3 && x >= xt && d_perp <= 0 &&
		    d_perp >= -thick) {
			*eta = 1e23; *rho = 3350.0;
		}
		if (inz && y > 5e3 && y <= 150e3 && x >= xt && d_perp > 0 &&
		    d_perp <= 10e3) {
			*eta = 1e17; *rho = 3250.0;
		}
		*t = 273.0 + (1600.0 - 273.0) * erf(y / (2.0 * sqrt(kappa * age)));
	}
}

int i3g_gen_config(int which, long nx, long ny, long nz, unsigned long long seed,
		   double *gx, double *gy, double *gz, float *nu, float *ro,
		   float *tk, float *nxx, float *nxy, float *nxz, float *nyz,
		   double *params)
{
	cfg_t c;
	long i, j, k;
	if (which < 0 || which > 2 || nx < 5 || ny < 5 || nz < 5) return I3G_EINVAL;
	memset(&c, 0, sizeof(c));
	c.which = which;
	if (which == 0) { c.xs = c.ys = c.zs = 100e3; }
	else if (which == 1) { c.xs = c.ys = c.zs = 500e3; }
	else { c.xs = 1000e3; c.ys = 500e3; c.zs = 1000e3; }
	if (which == 1) {
		c.nblob = 32;
		for (k = 0; k < 32; k++) {
			c.bx[k] = urand(seed, 1, k) * c.xs;
			c.by[k] = urand(seed, 2, k) * c.ys;
			c.bz[k] = urand(seed, 3, k) * c.zs;
			c.bs[k] = 20e3 + 40e3 * urand(seed, 4, k);
			c.ba[k] = -2.0 + 4.0 * urand(seed, 5, k);
		}
	}
	for (i = 0; i < nx; i++) gx[i] = c.xs * (double)i / (double)(nx - 1);
	for (j = 0; j < ny; j++) gy[j] = c.ys * (double)j / (double)(ny - 1);
	for (k = 0; k < nz; k++) gz[k] = c.zs * (double)k / (double)(nz - 1);
	if (params) {
		params[0] = which == 2 ? 1e17 : 1e18;
		params[1] = 1e25;
	}
#pragma omp parallel for private(i, j)
	for (k = 0; k < nz; k++)
	for (i = 0; i < nx; i++)
	for (j = 0; j < ny; j++) {
		long n = k * nx * ny + i * ny + j;
		double e, r, t;
		double xm = i > 0 ? 0.5 * (gx[i] + gx[i - 1]) : gx[0];
		double ym = j > 0 ? 0.5 * (gy[j] + gy[j - 1]) : gy[0];
		double zm = k > 0 ? 0.5 * (gz[k] + gz[k - 1]) : gz[0];
		double xp = i < nx - 1 ? 0.5 * (gx[i] + gx[i + 1]) : gx[nx - 1];
		double yp = j < ny - 1 ? 0.5 * (gy[j] + gy[j + 1]) : gy[ny - 1];
		double zp = k < nz - 1 ? 0.5 * (gz[k] + gz[k + 1]) : gz[nz - 1];
		material(&c, gx[i], gy[j], gz[k], &e, &r, &t);
		nu[n] = (float)e; ro[n] = (float)r;
		if (tk) tk[n] = (float)t;
		/* cell centre with max corner (i,j,k) */
		material(&c, xm, ym, zm, &e, &r, &t);
		nxx[n] = (float)e;
		/* z-edge (x[i], y[j], zmid[k]) */
		material(&c, gx[i], gy[j], zp, &e, &r, &t);
		nxy[n] = (float)e;
		/* y-edge (x[i], ymid[j], z[k]) */
		material(&c, gx[i], yp, gz[k], &e, &r, &t);
		nxz[n] = (float)e;
		/* x-edge (xmid[i], y[j], z[k]) */
		material(&c, xp, gy[j], gz[k], &e, &r, &t);
		nyz[n] = (float)e;
	}
	return I3G_OK;
}

/* Markers of the seeded configurations: per_axis^3 markers per cell at
 * jittered sub-cell centres (the in3mg.c:182-184 analogue: offset
 * U(-1/2, 1/2) of the marker spacing per axis), ids in cell order (z, x, y
 * cells, then the sub-cell index).  Rock type from the layout at the
 * marker: configuration 2 gives 2 asthenosphere, 3 lithosphere, 4 slab,
 * 5 weak channel; configurations 0 and 1 give 2 background, 3 inclusion /
 * positive blob anomaly.  Temperature from the layout's profile. */
long i3g_gen_markers(int which, long nx, long ny, long nz, unsigned long long seed, const double *gx,
		     const double *gy, const double *gz, int per_axis, float *x, float *y, float *z, float *tk,
		     signed char *type)
{
	cfg_t c;
	long k;
	if (which < 0 || which > 2 || nx < 2 || ny < 2 || nz < 2 || per_axis < 1 || !gx || !gy || !gz) return I3G_EINVAL;
	memset(&c, 0, sizeof(c));
	c.which = which;
	c.xs = gx[nx - 1]; c.ys = gy[ny - 1]; c.zs = gz[nz - 1];
	if (which == 1) {
		c.nblob = 32;
		for (k = 0; k < 32; k++) {
			c.bx[k] = urand(seed, 1, k) * c.xs;
			c.by[k] = urand(seed, 2, k) * c.ys;
			c.bz[k] = urand(seed, 3, k) * c.zs;
			c.bs[k] = 20e3 + 40e3 * urand(seed, 4, k);
			c.ba[k] = -2.0 + 4.0 * urand(seed, 5, k);
		}
	}
	long pc = (long)per_axis * per_axis * per_axis;
	long ncell = (nx - 1) * (ny - 1) * (nz - 1);
	if (!x) return ncell * pc;
	if (!y || !z || !tk || !type) return I3G_EINVAL;
#pragma omp parallel for
	for (k = 0; k < nz - 1; k++) {
		long i, j, q;
		for (i = 0; i < nx - 1; i++)
			for (j = 0; j < ny - 1; j++)
				for (q = 0; q < pc; q++) {
					long cell = (k * (nx - 1) + i) * (ny - 1) + j, id = cell * pc + q;
					int a = (int)(q / (per_axis * per_axis)), b = (int)((q / per_axis) % per_axis),
					    d = (int)(q % per_axis);
					double sx = (gx[i + 1] - gx[i]) / per_axis, sy = (gy[j + 1] - gy[j]) / per_axis,
					       sz = (gz[k + 1] - gz[k]) / per_axis;
					double px = gx[i] + (a + 0.5 + (urand(seed, 11, id) - 0.5)) * sx;
					double py = gy[j] + (d + 0.5 + (urand(seed, 12, id) - 0.5)) * sy;
					double pz = gz[k] + (b + 0.5 + (urand(seed, 13, id) - 0.5)) * sz;
					double e, r, t;
					material(&c, px, py, pz, &e, &r, &t);
					int ty = 2;
					if (which == 2) {
						if (e < 1e18) ty = 5;
						else if (e > 1e22) ty = r > 3325.0 ? 4 : 3;
					} else if (which == 0) {
						ty = e > 1e22 ? 3 : 2;
					} else {
						ty = r < 3300.0 ? 3 : 2;
					}
					x[id] = (float)px; y[id] = (float)py; z[id] = (float)pz;
					tk[id] = (float)t;
					type[id] = (signed char)ty;
				}
	}
	return ncell * pc;
}

const char *i3g_error_string(int code)
{
	switch (code) {
	case I3G_OK: return "ok";
	case I3G_EINVAL: return "invalid argument";
	case I3G_ENOMEM: return "out of memory";
	case I3G_EDEVICE: return "HIP runtime error";
	case I3G_ENOCONV: return "solver did not converge";
	case I3G_EBC: return "unsupported boundary-condition table";
	case I3G_ECAP: return "capacity exceeded";
	default: return "unknown error";
	}
}
