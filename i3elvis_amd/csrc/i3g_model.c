/*
 * i3g time-step loop: the per-step sequence of the reference's main loop
 * (i3mg.c:80-151: vpiterate, titerate, movemark, ronurecalc) in plain C
 * over the C ABI of include/i3g.h, with every node field resident on the
 * device.  The stages exchange device arrays; the host only sees scalars
 * (time step, residuals, timings).  Each ABI call synchronises its own
 * stream before it returns, so the wall-clock stage times are the stages'
 * own.
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/i3g.h"

#define MODEL_MAXROCK 64

/* node field planes of the marker -> node transfer (i3g_markers_to_nodes
 * order) */
enum { F_NU, F_RO, F_CP, F_KT, F_HT, F_ET, F_TK, F_NXX, F_NXY, F_NXZ, F_NYZ, F_COUNT };

struct i3g_model {
	long nx, ny, nz, N;
	i3g_ctx *st;
	i3g_heat *heat;
	i3g_markers *mk;
	float *fld;   /* F_COUNT x N node fields */
	float *tk0;   /* temperature at the start of the heat step */
	float *vel;   /* 3 x N velocities of the last solve */
	double *pr;   /* pressure of the last solve */
	double phys[6];
	i3g_rock rocks[MODEL_MAXROCK];
	int nrock;
	double timesum;
	int solved;   /* a previous solution is resident */
};

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int i3g_model_destroy(i3g_model *m)
{
	if (!m) return I3G_OK;
	if (m->mk) i3g_markers_destroy(m->mk);
	if (m->heat) i3g_heat_destroy(m->heat);
	if (m->st) i3g_destroy(m->st);
	if (m->fld) hipFree(m->fld);
	if (m->tk0) hipFree(m->tk0);
	if (m->vel) hipFree(m->vel);
	if (m->pr) hipFree(m->pr);
	free(m);
	return I3G_OK;
}

int i3g_model_create(const i3g_setup *s, long marker_capacity, i3g_model **out)
{
	int rc;
	if (!s || !out || marker_capacity < 1) return I3G_EINVAL;
	*out = NULL;
	i3g_model *m = (i3g_model *)calloc(1, sizeof(*m));
	if (!m) return I3G_ENOMEM;
	m->nx = s->nx; m->ny = s->ny; m->nz = s->nz;
	m->N = s->nx * s->ny * s->nz;
	m->phys[1] = 9.81; m->phys[3] = 1e18; m->phys[4] = 1e25;
	if ((rc = i3g_create(s, &m->st)) || (rc = i3g_heat_create(s, &m->heat)) ||
	    (rc = i3g_markers_create(s->nx, s->ny, s->nz, s->gx, s->gy, s->gz, marker_capacity, &m->mk))) {
		i3g_model_destroy(m);
		return rc;
	}
	/* tkrecalc after every transfer: the T block of the BC tables */
	rc = i3g_markers_set_tbc(m->mk, s->bondm + 8 * m->N, s->bondv, s->bondn, s->nrec);
	if (rc) { i3g_model_destroy(m); return rc; }
	long N = m->N;
	if (hipMalloc((void **)&m->fld, (size_t)F_COUNT * N * sizeof(float)) != hipSuccess ||
	    hipMalloc((void **)&m->tk0, (size_t)N * sizeof(float)) != hipSuccess ||
	    hipMalloc((void **)&m->vel, (size_t)3 * N * sizeof(float)) != hipSuccess ||
	    hipMalloc((void **)&m->pr, (size_t)N * sizeof(double)) != hipSuccess) {
		i3g_model_destroy(m);
		return I3G_ENOMEM;
	}
	if (hipMemset(m->fld, 0, (size_t)F_COUNT * N * sizeof(float)) != hipSuccess ||
	    hipMemset(m->vel, 0, (size_t)3 * N * sizeof(float)) != hipSuccess ||
	    hipMemset(m->pr, 0, (size_t)N * sizeof(double)) != hipSuccess) {
		i3g_model_destroy(m);
		return I3G_EDEVICE;
	}
	*out = m;
	return I3G_OK;
}

int i3g_model_set_markers(i3g_model *m, long n, const float *x, const float *y, const float *z,
			  const float *tk, const signed char *type, const i3g_rock *rocks, int nrock)
{
	if (!m || n < 0 || !x || !y || !z || !tk || !type || !rocks || nrock < 1 || nrock > MODEL_MAXROCK)
		return I3G_EINVAL;
	float *eta = (float *)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
	if (!eta) return I3G_ENOMEM;
	int rc = i3g_markers_upload(m->mk, n, x, y, z, tk, eta, type, NULL);
	free(eta);
	if (rc) return rc;
	memcpy(m->rocks, rocks, (size_t)nrock * sizeof(i3g_rock));
	m->nrock = nrock;
	return I3G_OK;
}

int i3g_model_set_physics(i3g_model *m, const double *phys)
{
	if (!m || !phys) return I3G_EINVAL;
	memcpy(m->phys, phys, sizeof(m->phys));
	return I3G_OK;
}

/* ronurecalc: cell re-sort and marker -> node transfer into the resident
 * node fields (the marker viscosity recomputed from its temperature and
 * pressure every time) */
static int transfer(i3g_model *m, const i3g_model_params *p, double *t_sort, double *t_transfer)
{
	long N = m->N;
	float *f = m->fld;
	double t0 = now_s();
	int rc = i3g_markers_sort(m->mk, NULL);
	if (rc) return rc;
	double t1 = now_s();
	i3g_mk_params mp = p->mk;
	mp.timestep = 0;
	rc = i3g_markers_to_nodes(m->mk, m->rocks, m->nrock, &mp, m->pr, f + F_NU * N, f + F_RO * N, f + F_CP * N,
				  f + F_KT * N, f + F_HT * N, f + F_ET * N, f + F_TK * N, f + F_NXX * N,
				  f + F_NXY * N, f + F_NXZ * N, f + F_NYZ * N);
	double t2 = now_s();
	if (t_sort) *t_sort = t1 - t0;
	if (t_transfer) *t_transfer = t2 - t1;
	return rc;
}

int i3g_model_init(i3g_model *m, const i3g_model_params *p, const double *pr)
{
	if (!m || !p || m->nrock < 1) return I3G_EINVAL;
	if (pr && hipMemcpy(m->pr, pr, (size_t)m->N * sizeof(double), hipMemcpyDefault) != hipSuccess)
		return I3G_EDEVICE;
	m->timesum = 0;
	m->solved = 0;
	return transfer(m, p, NULL, NULL);
}

int i3g_model_step(i3g_model *m, const i3g_model_params *p, i3g_net *net, i3g_model_stats *st)
{
	if (!m || !p || m->nrock < 1) return I3G_EINVAL;
	long N = m->N;
	float *f = m->fld;
	i3g_model_stats s;
	memset(&s, 0, sizeof(s));
	double t0 = now_s();
	/* vpiterate: coefficients from the last transfer, the start (previous
	 * solution, or the learned start with it as the fallback), the solve,
	 * the fields and the displacement-limited time step */
	int rc = i3g_set_coefficients(m->st, f + F_NU * N, f + F_NXX * N, f + F_NXY * N, f + F_NXZ * N,
				      f + F_NYZ * N, f + F_RO * N, m->phys);
	if (rc) return rc;
	if (p->learned && net) {
		long gauge[3] = {p->stokes.pxinit, p->stokes.pyinit, p->stokes.pzinit};
		double info[6];
		rc = i3g_stokes_learned_start_ex(m->st, net, p->pscale, p->vscale, gauge, m->solved ? 1 : 0, info);
		if (rc) return rc;
		s.learned_taken = info[5] != 0;
	}
	i3g_stats ss;
	rc = i3g_stokes_solve_resident(m->st, &p->stokes, &ss);
	if (rc && rc != I3G_ENOCONV) return rc;
	m->solved = 1;
	s.vcycles = ss.vcycles;
	s.outer = ss.outer_iters;
	s.rel_res = ss.rel_res;
	double vmax[3];
	rc = i3g_stokes_fields(m->st, m->vel, m->vel + N, m->vel + 2 * N, m->pr, vmax);
	if (rc) return rc;
	double dt = p->maxtmstep;
	for (int a = 0; a < 3; a++)
		if (vmax[a] > 0 && dt > p->maxxystep / vmax[a]) dt = p->maxxystep / vmax[a];
	s.timestep = dt;
	double t1 = now_s();
	/* titerate: heat step from the transferred temperature, then the
	 * markers take the nodal change */
	if (hipMemcpy(m->tk0, f + F_TK * N, (size_t)N * sizeof(float), hipMemcpyDeviceToDevice) != hipSuccess)
		return I3G_EDEVICE;
	i3g_heat_params hp = p->heat;
	hp.timestep = dt;
	hp.timesum = m->timesum;
	i3g_heat_stats hs;
	rc = i3g_heat_solve(m->heat, f + F_TK * N, f + F_RO * N, f + F_CP * N, f + F_KT * N, f + F_HT * N, &hp, &hs);
	if (rc && rc != I3G_ENOCONV) return rc;
	s.heat_substeps = hs.substeps;
	s.heat_sweeps = hs.sweeps;
	double t2 = now_s();
	if (hs.substeps > 0) {
		rc = i3g_markers_update_tk(m->mk, f + F_TK * N, m->tk0, f + F_KT * N, f + F_RO * N, f + F_CP * N,
					   p->heatdif, dt, 1e30, 0.0);
		if (rc) return rc;
	}
	double t3 = now_s();
	/* movemark: RK4 in the new velocity field */
	rc = i3g_markers_advect_rk4(m->mk, dt, m->vel, m->vel + N, m->vel + 2 * N);
	if (rc) return rc;
	double t4 = now_s();
	/* ronurecalc for the next step */
	rc = transfer(m, p, &s.t_sort, &s.t_transfer);
	if (rc) return rc;
	m->timesum += dt;
	s.timesum = m->timesum;
	s.t_stokes = t1 - t0;
	s.t_heat = t2 - t1;
	s.t_mktk = t3 - t2;
	s.t_rk4 = t4 - t3;
	s.seconds = now_s() - t0;
	if (st) *st = s;
	return I3G_OK;
}

int i3g_model_get(i3g_model *m, float *vx, float *vy, float *vz, double *pr, float *tk, float *nu)
{
	if (!m) return I3G_EINVAL;
	long N = m->N;
	size_t fb = (size_t)N * sizeof(float);
	if ((vx && hipMemcpy(vx, m->vel, fb, hipMemcpyDefault) != hipSuccess) ||
	    (vy && hipMemcpy(vy, m->vel + N, fb, hipMemcpyDefault) != hipSuccess) ||
	    (vz && hipMemcpy(vz, m->vel + 2 * N, fb, hipMemcpyDefault) != hipSuccess) ||
	    (pr && hipMemcpy(pr, m->pr, (size_t)N * sizeof(double), hipMemcpyDefault) != hipSuccess) ||
	    (tk && hipMemcpy(tk, m->fld + F_TK * N, fb, hipMemcpyDefault) != hipSuccess) ||
	    (nu && hipMemcpy(nu, m->fld + F_NU * N, fb, hipMemcpyDefault) != hipSuccess))
		return I3G_EDEVICE;
	return I3G_OK;
}

i3g_markers *i3g_model_markers(i3g_model *m) { return m ? m->mk : NULL; }
i3g_ctx *i3g_model_stokes(i3g_model *m) { return m ? m->st : NULL; }
