/*
 * Temperature boxes of the initial condition on the device (in3mg.c:
 * 806-951): every box in deck order sets the node temperatures inside it.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/i3g.h"

/* m1serch..m3serch (mark3mg.c:1021-1087): the grid line below v by
 * bisection, at most n - 2 */
__device__ inline int gen_serch(const double *g, int n, double v)
{
	int lo = 0, hi = n - 1;
	do {
		int mid = (lo + hi) / 2;
		if (g[mid] > v) hi = mid;
		else lo = mid;
	} while (hi - lo > 1);
	if (lo > n - 2) lo = n - 2;
	return lo;
}

/* tkcalc1 (mark3mg.c:2844-2910): the field interpolated at a point from
 * the eight nodes of its cell, summed in the reference's order */
__device__ inline double gen_tkcalc(const float *tk, const double *gx, const double *gy, const double *gz, int nx,
				    int ny, int nz, double x, double y, double z)
{
#pragma clang fp contract(off)
	int m10 = gen_serch(gx, nx, x), m20 = gen_serch(gy, ny, y), m30 = gen_serch(gz, nz, z);
	double dx = gx[m10 + 1] - gx[m10], dy = gy[m20 + 1] - gy[m20], dz = gz[m30 + 1] - gz[m30];
	double s = 0;
	for (int m3 = m30; m3 <= m30 + 1; m3++) {
		double ddz = 1.0 - fabs((gz[m3] - z) / dz);
		for (int m1 = m10; m1 <= m10 + 1; m1++) {
			double ddx = 1.0 - fabs((gx[m1] - x) / dx);
			for (int m2 = m20; m2 <= m20 + 1; m2++) {
				double ddy = 1.0 - fabs((gy[m2] - y) / dy);
				double swt = ddx * ddy * ddz;
				s += swt * tk[(long)m3 * nx * ny + (long)m1 * ny + m2];
			}
		}
	}
	return s;
}

struct TBox {
	int type;
	double bx[8], by[8], bz[8], t[8];
	double xs, ys, zs;
};

/* position of node (m1, m2, m3) in the box's bilinear frame
 * (in3mg.c:841-844); true when inside */
__device__ inline bool box_frame(const TBox &B, double gxv, double gyv, double gzv, double &x, double &y, double &z)
{
#pragma clang fp contract(off)
	const double *bx = B.bx, *by = B.by, *bz = B.bz;
	z = (gzv - bz[0]) / (bz[7] - bz[0]);
	x = (gxv - (bx[0] * (1.0 - z) + bx[5] * z)) / ((bx[2] - bx[0]) * (1.0 - z) + (bx[7] - bx[5]) * z);
	y = (gyv - (by[0] * (1.0 - x) * (1.0 - z) + by[2] * x * (1.0 - z) + by[4] * (1.0 - x) * z + by[6] * x * z)) /
	    ((by[1] - by[0]) * (1.0 - x) * (1.0 - z) + (by[3] - by[2]) * x * (1.0 - z) + (by[5] - by[4]) * (1.0 - x) * z +
	     (by[7] - by[6]) * x * z);
	return x >= 0 && x <= 1.0 && y >= 0 && y <= 1.0 && z >= 0 && z <= 1.0;
}

/* types 0, 1, 4, 2, 3: one thread per node */
__global__ void k_tbox_node(TBox B, const double *gx, const double *gy, const double *gz, int nx, int ny, int nz,
			    float *tk)
{
#pragma clang fp contract(off)
	long N = (long)nx * ny * nz;
	long m4 = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (m4 >= N) return;
	int m3 = (int)(m4 / ((long)nx * ny));
	long r = m4 - (long)m3 * nx * ny;
	int m1 = (int)(r / ny), m2 = (int)(r - (long)m1 * ny);
	const double *bx = B.bx, *by = B.by, *bz = B.bz, *t = B.t;
	if (B.type != 2 && B.type != 3) {
		double x, y, z;
		if (!box_frame(B, gx[m1], gy[m2], gz[m3], x, y, z)) return;
		if (B.type == 0) {
			/* the reference accumulates into the float node value */
			float v = (float)((1.0 - x) * (1.0 - y) * (1.0 - z) * t[0]);
			v = (float)(v + (1.0 - x) * (y) * (1.0 - z) * t[1]);
			v = (float)(v + (x) * (1.0 - y) * (1.0 - z) * t[2]);
			v = (float)(v + (x) * (y) * (1.0 - z) * t[3]);
			v = (float)(v + (1.0 - x) * (1.0 - y) * (z)*t[4]);
			v = (float)(v + (1.0 - x) * (y) * (z)*t[5]);
			v = (float)(v + (x) * (1.0 - y) * (z)*t[6]);
			v = (float)(v + (x) * (y) * (z)*t[7]);
			tk[m4] = v;
		} else if (B.type == 1 || B.type == 4) {
			double ival = t[2] * (1.0 - x) * (1.0 - z) + t[3] * x * (1.0 - z) + t[4] * (1.0 - x) * z + t[5] * x * z;
			ival = (gy[m2] - (by[0] * (1.0 - x) * (1.0 - z) + by[2] * x * (1.0 - z) + by[4] * (1.0 - x) * z +
					  by[6] * x * z)) /
			       2.0 / pow(t[6] * ival * 3.15576e+7, 0.5);
			tk[m4] = (float)(t[1] - (1.0 - erf(ival)) * (t[1] - t[0]) + t[7] * (gy[m2] - by[0]));
		}
		return;
	}
	double dx = (gx[m1] - bx[0]) / bx[7];
	double dy = (gy[m2] - by[0]) / by[7];
	double dz = (gz[m3] - bz[0]) / bz[7];
	double ival;
	bool in;
	if (B.type == 2) {
		ival = pow(dx * dx + dy * dy + dz * dz, 0.5);
		in = ival >= bx[2] / B.xs && ival <= bx[5] / B.xs;
	} else {
		ival = pow(dx * dx + dz * dz, 0.5);
		in = ival >= bx[2] / B.xs && ival <= bx[5] / B.xs && fabs(dy) <= 1.0;
	}
	if (in) tk[m4] = (float)(t[0] + (ival - bx[2] / B.xs) / (bx[5] / B.xs - bx[2] / B.xs) * (t[1] - t[0]));
}

/* types 5, 6: a node reads the field at the box's corner columns at its own
 * depth, where nodes of the same y line may already hold this box's
 * values; the reference's loop (m3, m1, m2) is kept by one thread per y
 * line walking m3 then m1 (its interpolation gives the other y line the
 * weight 0) */
__global__ void k_tbox_transition(TBox B, const double *gx, const double *gy, const double *gz, int nx, int ny, int nz,
				  float *tk)
{
#pragma clang fp contract(off)
	int m2 = blockIdx.x * blockDim.x + threadIdx.x;
	if (m2 >= ny) return;
	const double *bx = B.bx, *bz = B.bz;
	double gyv = gy[m2];
	for (int m3 = 0; m3 < nz; m3++)
		for (int m1 = 0; m1 < nx; m1++) {
			double x, y, z;
			if (!box_frame(B, gx[m1], gyv, gz[m3], x, y, z)) continue;
			double t0 = gen_tkcalc(tk, gx, gy, gz, nx, ny, nz, bx[0], gyv, bz[0]);
			double t1 = gen_tkcalc(tk, gx, gy, gz, nx, ny, nz, bx[2], gyv, bz[0]);
			double t2 = gen_tkcalc(tk, gx, gy, gz, nx, ny, nz, bx[5], gyv, bz[7]);
			double t3 = gen_tkcalc(tk, gx, gy, gz, nx, ny, nz, bx[7], gyv, bz[7]);
			tk[(long)m3 * nx * ny + (long)m1 * ny + m2] =
				(float)(t0 * (1.0 - x) * (1.0 - z) + t1 * x * (1.0 - z) + t2 * (1.0 - x) * z + t3 * x * z);
		}
}

extern "C" int i3g_temperature_boxes(long nx, long ny, long nz, const double *gx, const double *gy, const double *gz,
				     double xsize, double ysize, double zsize, const i3g_temp_box *boxes, int nbox,
				     float *tk)
{
	if (nx < 2 || ny < 2 || nz < 2 || !gx || !gy || !gz || !boxes || nbox < 0 || !tk) return I3G_EINVAL;
	for (int b = 0; b < nbox; b++) {
		int ty = boxes[b].type;
		if (ty < 0 || ty > 6) return I3G_EINVAL;
	}
	long N = nx * ny * nz;
	double *dg = nullptr;
	float *dt = nullptr;
	if (hipMalloc((void **)&dg, (nx + ny + nz) * sizeof(double)) != hipSuccess) return I3G_ENOMEM;
	if (hipMalloc((void **)&dt, N * sizeof(float)) != hipSuccess) {
		hipFree(dg);
		return I3G_ENOMEM;
	}
	hipError_t e = hipMemcpy(dg, gx, nx * sizeof(double), hipMemcpyDefault);
	if (e == hipSuccess) e = hipMemcpy(dg + nx, gy, ny * sizeof(double), hipMemcpyDefault);
	if (e == hipSuccess) e = hipMemcpy(dg + nx + ny, gz, nz * sizeof(double), hipMemcpyDefault);
	if (e == hipSuccess) e = hipMemcpy(dt, tk, N * sizeof(float), hipMemcpyDefault);
	for (int b = 0; b < nbox && e == hipSuccess; b++) {
		const i3g_temp_box &r = boxes[b];
		TBox B;
		B.type = r.type;
		for (int q = 0; q < 8; q++) B.bx[q] = B.by[q] = B.bz[q] = 0;
		B.bx[0] = r.x0; B.bx[2] = r.x2; B.by[0] = r.y0; B.by[1] = r.y1; B.by[2] = r.y2; B.by[3] = r.y3;
		B.bz[0] = r.z0; B.bx[5] = r.x5; B.bx[7] = r.x7; B.by[4] = r.y4; B.by[5] = r.y5; B.by[6] = r.y6;
		B.by[7] = r.y7; B.bz[7] = r.z7;
		for (int q = 0; q < 8; q++) B.t[q] = r.t[q];
		B.xs = xsize; B.ys = ysize; B.zs = zsize;
		if (r.type == 5 || r.type == 6)
			hipLaunchKernelGGL(k_tbox_transition, dim3((unsigned)((ny + 63) / 64)), dim3(64), 0, 0, B, (const double *)dg,
					   (const double *)(dg + nx), (const double *)(dg + nx + ny), (int)nx, (int)ny, (int)nz, dt);
		else
			hipLaunchKernelGGL(k_tbox_node, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, B, (const double *)dg,
					   (const double *)(dg + nx), (const double *)(dg + nx + ny), (int)nx, (int)ny, (int)nz, dt);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipMemcpy(tk, dt, N * sizeof(float), hipMemcpyDefault);
	hipFree(dg);
	hipFree(dt);
	return e == hipSuccess ? I3G_OK : I3G_EDEVICE;
}
