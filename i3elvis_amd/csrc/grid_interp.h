/*
 * Device point interpolation on the level-0 grid shared by the marker and
 * heat paths: the reference's line search (m1serch..m3serch) and its
 * staggered trilinear velocity (vxyzcalc1).  Products and sums are rounded
 * one at a time, as the reference's x86-64 build.
 */
#pragma once
#include <hip/hip_runtime.h>

/* largest i with g[i] <= v (binary search as m1serch, mark3mg.c:1018-1037),
 * clamped to n-2 */
__device__ __forceinline__ int serch(const double *g, int n, double v)
{
#pragma clang fp contract(off)
	int lo = 0, hi = n - 1, mid;
	do {
		mid = (lo + hi) / 2;
		if (g[mid] > v) hi = mid; else lo = mid;
	} while (hi - lo > 1);
	if (lo > n - 2) lo = n - 2;
	return lo;
}

/* The same index found by walking from a nearby start i (0 <= i <= n-2),
 * for points close to the previous one: upward while g[i+1] is not above
 * v, then downward while g[i] is above v.  For increasing lines this ends
 * at the largest i <= n-2 with g[i] <= v (0 below g[0]) and, as the
 * bisection, at n-2 for a NaN v. */
__device__ __forceinline__ int serch_from(const double *g, int n, double v, int i)
{
	while (i < n - 2 && !(g[i + 1] > v)) i++;
	while (i > 0 && g[i] > v) i--;
	return i;
}

/* staggered trilinear velocity at a point, vxyzcalc1 (mark3mg.c:2347-2530);
 * cell[3]: the node-line indices of the point, returned, and when
 * from_hint is set also the start of the search (the previous point's) */
__device__ inline void vxyz(const float *vx, const float *vy, const float *vz, const double *gx, const double *gy,
			    const double *gz, int nx, int ny, int nz, double x, double y, double z, double *v,
			    int *cell = nullptr, bool from_hint = false)
{
#pragma clang fp contract(off)
	double xs = gx[nx - 1], ys = gy[ny - 1], zs = gz[nz - 1];
	if (x < 0) x = 0;
	if (x > xs) x = xs;
	if (y < 0) y = 0;
	if (y > ys) y = ys;
	if (z < 0) z = 0;
	if (z > zs) z = zs;
	int m1, m2, m3;
	if (cell && from_hint) {
		m1 = serch_from(gx, nx, x, cell[0]);
		m2 = serch_from(gy, ny, y, cell[1]);
		m3 = serch_from(gz, nz, z, cell[2]);
	} else {
		m1 = serch(gx, nx, x);
		m2 = serch(gy, ny, y);
		m3 = serch(gz, nz, z);
	}
	if (cell) {
		cell[0] = m1;
		cell[1] = m2;
		cell[2] = m3;
	}
	long sx = ny, sz = (long)nx * ny;
	for (int comp = 0; comp < 3; comp++) {
		int a = m1, b = m2, c = m3;
		/* node-aligned axis: comp; the other two use cell-mid lines */
		if (comp == 0) {
			if (a > nx - 2) a = nx - 2;
		} else {
			if (x < (gx[a] + gx[a + 1]) / 2.0) a -= 1;
			if (a < 0) a = 0;
			if (a > nx - 3) a = nx - 3;
		}
		if (comp == 1) {
			if (b > ny - 2) b = ny - 2;
		} else {
			if (y < (gy[b] + gy[b + 1]) / 2.0) b -= 1;
			if (b < 0) b = 0;
			if (b > ny - 3) b = ny - 3;
		}
		if (comp == 2) {
			if (c > nz - 2) c = nz - 2;
		} else {
			if (z < (gz[c] + gz[c + 1]) / 2.0) c -= 1;
			if (c < 0) c = 0;
			if (c > nz - 3) c = nz - 3;
		}
		double dx = comp == 0 ? gx[a + 1] - gx[a] : (gx[a + 2] - gx[a]) / 2.0;
		double dy = comp == 1 ? gy[b + 1] - gy[b] : (gy[b + 2] - gy[b]) / 2.0;
		double dz = comp == 2 ? gz[c + 1] - gz[c] : (gz[c + 2] - gz[c]) / 2.0;
		const float *f = comp == 0 ? vx : comp == 1 ? vy : vz;
		double acc = 0;
		for (int m30 = c; m30 <= c + 1; m30++) {
			double ddz = comp == 2 ? (gz[m30] - z) / dz : ((gz[m30] + gz[m30 + 1]) / 2.0 - z) / dz;
			if (m30 == c) ddz = -ddz;
			ddz = 1.0 - ddz;
			for (int m10 = a; m10 <= a + 1; m10++) {
				double ddx = comp == 0 ? (gx[m10] - x) / dx : ((gx[m10] + gx[m10 + 1]) / 2.0 - x) / dx;
				if (m10 == a) ddx = -ddx;
				ddx = 1.0 - ddx;
				for (int m20 = b; m20 <= b + 1; m20++) {
					double ddy = comp == 1 ? (gy[m20] - y) / dy : ((gy[m20] + gy[m20 + 1]) / 2.0 - y) / dy;
					if (m20 == b) ddy = -ddy;
					ddy = 1.0 - ddy;
					acc += f[m30 * sz + m10 * sx + m20] * (ddx * ddy * ddz);
				}
			}
		}
		v[comp] = acc;
	}
}

