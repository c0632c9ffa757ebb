/*
 * i3g_glue.c -- drop-in glue between a program built on the reference's
 * global state (head3mg.c) and the i3g C ABI (include/i3g.h).
 *
 * Use: compile this file in the same translation unit as the program's
 * globals, after head3mg.c (the reference builds i3mg.c as one unit that
 * #includes its other sources), link libi3g.so, and call
 *     i3g_vpiterate(m0, f0)  where the program calls vpiterate(m0, f0)
 *     i3g_titerate(m0)       where it calls titerate(m0)
 *     i3g_ronurecalc()       where it calls ronurecalc()
 *     i3g_movemark()         where it calls movemark()
 * (i3mg.c:98-137).  The functions read and write the same globals as the
 * routines they replace; device state is created on first use and kept
 * between steps (re-created when the grid changes).
 *
 * What each replaces:
 *   i3g_vpiterate  the Stokes solve of vpiterate (move3mg.c:1585-2636):
 *                  solve to i3g_glue_rtol (relative FP64 residual) from the
 *                  current vx/vy/vz/pr, pressure gauge (pxinit..pinit),
 *                  the stress/strain-rate fields (value mode of
 *                  sxxcalc..syzcalc, on the device) and the marker
 *                  time-step limit (move3mg.c:2609-2632).
 *   i3g_titerate   the heat step of titerate (heat3mg.c:287-588) with the
 *                  adiabyn/frictyn heating terms, then the marker
 *                  temperature update (heat3mg.c:592-782).
 *   i3g_ronurecalc ronurecalc (mark3mg.c:404-1015) with the whole viscalc,
 *                  retyping, water/air mixing, phase densities and the
 *                  partially molten rocks of the melt types (21-49, from
 *                  markw and markex), with tkrecalc; densimod >= 2 stops
 *                  the program with a message.
 *   i3g_movemark   movemark (mark3mg.c:2-397): topography of the markers,
 *                  viscosity, strain and grain state at the old positions
 *                  (epscalc1 + viscalc), the motion (outgrid, markmod,
 *                  fluid and immobile markers), then the immobile-marker
 *                  regeneration and the refill of markers that left.
 * Errors print a message and exit, as the reference does on fatal
 * conditions.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "i3g.h"

double i3g_glue_rtol = 1e-8;
int i3g_glue_verbose = 0;
/* node averages of ronurecalc and titerate's marker update: 0 the
 * reference's serial marker order (its float sums to the bit), 1 the
 * library's cell-segmented FP64 sums (i3g_markers_set_order) */
int i3g_glue_order = 0;

static struct {
	i3g_ctx *ctx;
	i3g_heat *heat;
	i3g_markers *mk;
	long nx, ny, nz, mcap, nrec;
	int multinum;
	double gsum;
	double *P, *X, *Y, *Z;
	float *tkstart;
} G;

static void glue_fail(const char *what, int rc)
{
	fprintf(stderr, "EXIT PROGRAM: i3g %s failed: %s (%d)\n", what, i3g_error_string(rc), rc);
	exit(1);
}

/* grid lines and BC table size identify the grid; a change re-creates the
 * device state */
static double grid_sum(void)
{
	double s = 0;
	long i;
	for (i = 0; i < xnumx; i++) s += gx[i] * (i + 1);
	for (i = 0; i < ynumy; i++) s += gy[i] * (i + 2);
	for (i = 0; i < znumz; i++) s += gz[i] * (i + 3);
	return s;
}

static void glue_grid(void)
{
	int rc;
	double gs = grid_sum();
	if (G.ctx && G.nx == xnumx && G.ny == ynumy && G.nz == znumz && G.multinum == multinum &&
	    G.nrec == bondnum + 1 && G.gsum == gs)
		return;
	if (G.ctx) i3g_destroy(G.ctx);
	if (G.heat) i3g_heat_destroy(G.heat);
	if (G.mk) i3g_markers_destroy(G.mk);
	free(G.P); free(G.X); free(G.Y); free(G.Z); free(G.tkstart);
	G.ctx = NULL; G.heat = NULL; G.mk = NULL; G.mcap = 0;
	i3g_setup s;
	s.nx = xnumx; s.ny = ynumy; s.nz = znumz;
	s.gx = gx; s.gy = gy; s.gz = gz;
	s.multinum = multinum;
	s.bondm = bondm;
	s.bondv = &bondv1[0][0];
	s.bondn = bondn1;
	s.nrec = bondnum + 1;
	if ((rc = i3g_create(&s, &G.ctx))) glue_fail("create", rc);
	if ((rc = i3g_heat_create(&s, &G.heat))) glue_fail("heat_create", rc);
	G.nx = xnumx; G.ny = ynumy; G.nz = znumz; G.multinum = multinum; G.nrec = bondnum + 1; G.gsum = gs;
	long N = xnumx * ynumy * znumz;
	G.P = malloc(sizeof(double) * N); G.X = malloc(sizeof(double) * N);
	G.Y = malloc(sizeof(double) * N); G.Z = malloc(sizeof(double) * N);
	G.tkstart = malloc(sizeof(float) * N);
	if (!G.P || !G.X || !G.Y || !G.Z || !G.tkstart) glue_fail("host allocation", I3G_ENOMEM);
}

/* markers of the program, uploaded in their order (ids = positions) */
static void glue_markers(void)
{
	int rc;
	if (!G.mk || G.mcap < marknum) {
		if (G.mk) i3g_markers_destroy(G.mk);
		G.mcap = marknum + marknum / 4 + 1024;
		if ((rc = i3g_markers_create(xnumx, ynumy, znumz, gx, gy, gz, G.mcap, &G.mk))) glue_fail("markers_create", rc);
		if ((rc = i3g_markers_set_order(G.mk, i3g_glue_order))) glue_fail("markers_set_order", rc);
	}
	if ((rc = i3g_markers_upload(G.mk, marknum, markx, marky, markz, markk, markv, (const signed char *)markt, NULL)))
		glue_fail("markers_upload", rc);
	/* the rheology's per-marker state: strain, grain roughness,
	 * dislocation share */
	if (marknum && ((rc = i3g_markers_set_strain(G.mk, marke)) || (rc = i3g_markers_set_grain(G.mk, markrr, markrn))))
		glue_fail("markers_set_state", rc);
	/* water content and extracted melt, by id, for the melt types */
	if ((rc = i3g_markers_set_melt(G.mk, marknum, markw, markex))) glue_fail("markers_set_melt", rc);
}

/* device markers back into the program's arrays (by id): positions,
 * temperature, viscosity, type (ronurecalc's retyping, movemark's depth
 * rules), strain, grain roughness and dislocation share */
static void glue_markers_back(void)
{
	int rc;
	long n = marknum, i;
	float *buf = malloc(sizeof(float) * 8 * (n ? n : 1));
	signed char *t = malloc(n ? n : 1);
	int *id = malloc(sizeof(int) * (n ? n : 1));
	if (!buf || !t || !id) glue_fail("host allocation", I3G_ENOMEM);
	if ((rc = i3g_markers_download(G.mk, buf, buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n, t, id)))
		glue_fail("markers_download", rc);
	if (n && ((rc = i3g_markers_get_strain(G.mk, buf + 5 * n)) ||
		  (rc = i3g_markers_get_grain(G.mk, buf + 6 * n, buf + 7 * n))))
		glue_fail("markers_get_state", rc);
	for (i = 0; i < n; i++) {
		long j = id[i];
		markx[j] = buf[i]; marky[j] = buf[n + i]; markz[j] = buf[2 * n + i];
		markk[j] = buf[3 * n + i]; markv[j] = buf[4 * n + i];
		marke[j] = buf[5 * n + i]; markrr[j] = buf[6 * n + i]; markrn[j] = buf[7 * n + i];
		markt[j] = t[i];
	}
	free(buf); free(t); free(id);
}

/* the rock table of the program (marknu.. markht) for the device */
static int glue_rocks(i3g_rock *rocks)
{
	int t, nr = rocknum > 0 && rocknum <= 64 ? rocknum : 64;
	for (t = 0; t < nr; t++) {
		rocks[t].nu0 = marknu[t]; rocks[t].dh = markdh[t]; rocks[t].dv = markdv[t];
		rocks[t].n0 = markn0[t]; rocks[t].n1 = markn1[t];
		rocks[t].ro = markro[t]; rocks[t].bro = markbb[t]; rocks[t].aro = markaa[t];
		rocks[t].cp = markcp[t]; rocks[t].kt = markkt[t]; rocks[t].kf = markkf[t]; rocks[t].kp = markkp[t];
		rocks[t].ht = markht[t];
		rocks[t].ss = markss[t]; rocks[t].mm = markmm[t]; rocks[t].ll = markll[t]; rocks[t].s1 = marks1[t];
		rocks[t].a0 = marka0[t]; rocks[t].a1 = marka1[t]; rocks[t].b0 = markb0[t]; rocks[t].b1 = markb1[t];
		rocks[t].e0 = marke0[t]; rocks[t].e1 = marke1[t];
	}
	return nr;
}

/* i3g_mk_params from the program's globals, whole rheology */
static void glue_mk_params(i3g_mk_params *mp, int sedtype)
{
	memset(mp, 0, sizeof(*mp));
	mp->nubeg = nubeg; mp->nuend = nuend; mp->nukoef = nukoef; mp->densimod = densimod; mp->timestep = timestep;
	mp->vdeep = vdeep; mp->nudeep = nudeep; mp->viscmod = viscmod;
	mp->rheology = 1;
	mp->eroslev = eroslev; mp->sedilev = sedilev; mp->waterlev = waterlev; mp->sedtype = sedtype;
	mp->mnumy = mnumy > 0 ? mnumy : 1;
	mp->teclmin = teclmin; mp->teclmax = teclmax;
	mp->timesum = timesum; mp->convend = convend;
	mp->outgrid = outgrid;
	mp->maxpmelt = maxpmelt;
}

void i3g_vpiterate(int m0, int f0)
{
	long N, n, m1, m2, m3;
	int rc;
	(void)m0; (void)f0;
	if (xnumx * ynumy * znumz > MAXNNN) glue_fail("grid above the program's stress arrays", I3G_ECAP);
	glue_grid();
	N = xnumx * ynumy * znumz;
	double phys[6] = {GXKOEF, GYKOEF, GZKOEF, nubeg, nuend, (double)viscmod};
	if ((rc = i3g_set_coefficients(G.ctx, nu, nxx, nxy, nxz, nyz, ro, phys))) glue_fail("set_coefficients", rc);
	for (n = 0; n < N; n++) {
		G.P[n] = pr[n]; G.X[n] = vx[n]; G.Y[n] = vy[n]; G.Z[n] = vz[n];
	}
	i3g_stokes_params p;
	i3g_stats st;
	i3g_stokes_default_params(&p);
	p.rtol = i3g_glue_rtol;
	p.pxinit = (long)pxinit; p.pyinit = (long)pyinit; p.pzinit = (long)pzinit; p.pinit = pinit;
	p.verbose = i3g_glue_verbose;
	rc = i3g_stokes_solve(G.ctx, G.P, G.X, G.Y, G.Z, &p, &st);
	if (rc && rc != I3G_ENOCONV) glue_fail("stokes_solve", rc);
	if (rc == I3G_ENOCONV) fprintf(stderr, "i3g: Stokes solve stopped at relative residual %.3e\n", st.rel_res);
	double mnx = 1e300, mxx = -1e300, mny = 1e300, mxy = -1e300, mnz = 1e300, mxz = -1e300;
	for (n = 0; n < N; n++) {
		vx[n] = (float)G.X[n]; vy[n] = (float)G.Y[n]; vz[n] = (float)G.Z[n]; pr[n] = G.P[n];
	}
	/* value ranges over the existing velocity nodes, then the marker time
	 * step limit maxxystep / |v|max (move3mg.c:2609-2632) */
	for (m3 = 0; m3 < znumz; m3++)
		for (m1 = 0; m1 < xnumx; m1++)
			for (m2 = 0; m2 < ynumy; m2++) {
				n = m3 * xnumx * ynumy + m1 * ynumy + m2;
				if (m2 < ynumy - 1 && m3 < znumz - 1) {
					if (vx[n] < mnx) mnx = vx[n];
					if (vx[n] > mxx) mxx = vx[n];
				}
				if (m1 < xnumx - 1 && m3 < znumz - 1) {
					if (vy[n] < mny) mny = vy[n];
					if (vy[n] > mxy) mxy = vy[n];
				}
				if (m1 < xnumx - 1 && m2 < ynumy - 1) {
					if (vz[n] < mnz) mnz = vz[n];
					if (vz[n] > mxz) mxz = vz[n];
				}
			}
	double vm[3] = {mxx > -mnx ? mxx : -mnx, mxy > -mny ? mxy : -mny, mxz > -mnz ? mxz : -mnz};
	for (int a = 0; a < 3; a++)
		if (vm[a] > 0 && timestep > maxxystep / vm[a]) timestep = maxxystep / vm[a];
	/* stress and strain-rate fields of the solution (value mode of
	 * sxxcalc..syzcalc, computed on the device) */
	rc = i3g_stokes_strain_stress(G.ctx, exx, eyy, ezz, exy, exz, eyz, sxx, syy, szz, sxy, sxz, syz);
	if (rc) glue_fail("strain_stress", rc);
}

void i3g_titerate(int m0)
{
	long N, n;
	int rc;
	(void)m0;
	glue_grid();
	N = xnumx * ynumy * znumz;
	for (n = 0; n < N; n++) G.tkstart[n] = tk[n];
	i3g_heat_params hp;
	i3g_heat_stats hs;
	i3g_heat_default_params(&hp);
	hp.timestep = timestep;
	hp.maxtkstep = maxtkstep;
	hp.heatmin = HEATMIN;
	hp.t0koef = t0koef;
	hp.sweeps = multittt[0][0] > 0 ? multittt[0][0] : 1;
	hp.timesum = timesum; hp.timebeg = timebeg; hp.timeend = timeend; hp.tempend = tempend;
	hp.verbose = i3g_glue_verbose;
	{
		/* heaterr's adiabatic and shear heating from the current
		 * velocity, pressure and stress/strain-rate fields */
		double gk[3] = {GXKOEF, GYKOEF, GZKOEF};
		const float *sig[6] = {sxx, syy, szz, sxy, sxz, syz};
		const float *eps[6] = {exx, eyy, ezz, exy, exz, eyz};
		rc = i3g_heat_set_sources(G.heat, adiabyn, frictyn, gk, ysize, stp100, vx, vy, vz, pr, et, sig, eps);
		if (rc) glue_fail("heat_set_sources", rc);
	}
	rc = i3g_heat_solve(G.heat, tk, ro, cp, kt, ht, &hp, &hs);
	if (rc && rc != I3G_ENOCONV) glue_fail("heat_solve", rc);
	if (hs.substeps == 0) return; /* zero step: the reference returns before the marker update */
	glue_markers();
	if ((rc = i3g_markers_sort(G.mk, NULL))) glue_fail("markers_sort", rc);
	rc = i3g_markers_update_tk(G.mk, tk, G.tkstart, kt, ro, cp, heatdif, timestep, zdeep, tdeep);
	if (rc) glue_fail("markers_update_tk", rc);
	glue_markers_back();
}

void i3g_ronurecalc(void)
{
	int rc, t;
	glue_grid();
	glue_markers();
	if ((rc = i3g_markers_sort(G.mk, NULL))) glue_fail("markers_sort", rc);
	if ((rc = i3g_markers_set_tbc(G.mk, bondm + nodenum8, &bondv1[0][0], bondn1, bondnum + 1)))
		glue_fail("markers_set_tbc", rc);
	static i3g_rock rocks[64];
	int nr = glue_rocks(rocks);
	(void)t;
	/* the sediment type of this call, as ronurecalc derives it
	 * (mark3mg.c:421-424) */
	sedimnum++;
	long c1 = (long)(sedimnum / sedimcyc), c2 = (c1 / 2) * 2;
	i3g_mk_params mp;
	glue_mk_params(&mp, c2 == c1 ? 3 : 4);
	/* viscalc's strain rates and stresses (epscalc1 at the markers) */
	const float *st12[12] = {exx, eyy, ezz, exy, exz, eyz, sxx, syy, szz, sxy, sxz, syz};
	if ((rc = i3g_markers_set_stress(G.mk, st12))) glue_fail("markers_set_stress", rc);
	rc = i3g_markers_to_nodes(G.mk, rocks, nr, &mp, pr, nu, ro, cp, kt, ht, et, tk, nxx, nxy, nxz, nyz);
	if (rc) glue_fail("markers_to_nodes", rc);
	glue_markers_back();
}

/* topography surface where the node density crosses topodens along y
 * (mark3mg.c:27-41), and each marker of types 2-19 with T > 0 that moves
 * gets its height (:75-88) */
static void glue_topography(void)
{
	long m1, m2, m3, m4, mm1;
	const double topodens = 2000.0;
	double *top = bufv + 2 * xnumx * znumz;
	for (m3 = 0; m3 < znumz; m3++)
		for (m1 = 0; m1 < xnumx; m1++)
			for (m2 = 1; m2 < ynumy; m2++) {
				m4 = m3 * xnumx * ynumy + m1 * ynumy + m2;
				if (ro[m4] > topodens && ro[m4 - 1] < topodens) {
					top[m3 * xnumx + m1] = gy[m2 - 1] + (gy[m2] - gy[m2 - 1]) * (topodens - ro[m4 - 1]) / (ro[m4] - ro[m4 - 1]);
					break;
				}
			}
	for (mm1 = 0; mm1 < marknum; mm1++) {
		int inside = markx[mm1] >= 0 && marky[mm1] >= 0 && markz[mm1] >= 0 && (double)markx[mm1] <= xsize &&
			     (double)marky[mm1] <= ysize && (double)markz[mm1] <= zsize;
		if (!(inside || (outgrid != 1 && (markt[mm1] < 50 || markt[mm1] >= 100)))) continue;
		if (!(markt[mm1] > 1 && markt[mm1] < 20 && markk[mm1] > 0)) continue;
		double x = markx[mm1], z = markz[mm1];
		if (x < 0) x = 0; if (x > xsize) x = xsize;
		if (z < 0) z = 0; if (z > zsize) z = zsize;
		long w0 = m1serch(x), w2 = m3serch(z);
		m4 = w2 * xnumx + w0;
		double dx = (markx[mm1] - gx[w0]) / (gx[w0 + 1] - gx[w0]);
		double dz = (markz[mm1] - gz[w2]) / (gz[w2 + 1] - gz[w2]);
		markto[mm1] = (float)((1.0 - dx) * (1.0 - dz) * top[m4] + dx * (1.0 - dz) * top[m4 + 1] +
				      (1.0 - dx) * dz * top[m4 + xnumx] + dx * dz * top[m4 + xnumx + 1]);
	}
}

/* the end of movemark on the host (mark3mg.c:290-385): when an immobile
 * marker (type >= 100) has moved farther than stp100 from its origin
 * (markk, markd, markw), every immobile marker inside the grid spawns a
 * normal marker of type - 100 and returns to its origin; spawned markers
 * then refill the slots of normal markers that left the grid */
static void glue_immobile(void)
{
	long mm1, marknum1 = marknum;
	int ynreset = 0;
	for (mm1 = 0; mm1 < marknum && !ynreset; mm1++)
		if (markt[mm1] >= 100) {
			double dx = markx[mm1] - markk[mm1], dy = marky[mm1] - markd[mm1], dz = markz[mm1] - markw[mm1];
			if (pow(dx * dx + dy * dy + dz * dz, 0.5) > stp100) ynreset = 1;
		}
	if (!ynreset) return;
	for (mm1 = 0; mm1 < marknum; mm1++)
		if (markt[mm1] >= 100) {
			if (markx[mm1] > 0 && marky[mm1] > 0 && markz[mm1] > 0 && (double)markx[mm1] < xsize &&
			    (double)marky[mm1] < ysize && (double)markz[mm1] < zsize) {
				if (marknum1 >= MAXMRK) glue_fail("marker arrays full", I3G_ECAP);
				markt[marknum1] = markt[mm1] - 100;
				markx[marknum1] = markx[mm1]; marky[marknum1] = marky[mm1]; markz[marknum1] = markz[mm1];
				markk[marknum1] = 0; markv[marknum1] = 0; markft[marknum1] = 0;
				markex[marknum1] = markex[mm1]; marke[marknum1] = marke[mm1];
				markd[marknum1] = -1.0; markw[marknum1] = -1.0; markrr[marknum1] = 0;
				marknum1++;
			}
			markx[mm1] = markk[mm1]; marky[mm1] = markd[mm1]; markz[mm1] = markw[mm1];
		}
	mm1 = 0;
	while (marknum1 > marknum && mm1 < marknum) {
		if (markt[mm1] < 100 && (markx[mm1] < 0 || marky[mm1] < 0 || markz[mm1] < 0 || (double)markx[mm1] > xsize ||
					 (double)marky[mm1] > ysize || (double)markz[mm1] > zsize)) {
			marknum1--;
			markt[mm1] = markt[marknum1];
			markk[mm1] = 0; markv[mm1] = 0; markft[mm1] = 0;
			markex[mm1] = markex[marknum1]; marke[mm1] = marke[marknum1];
			markd[mm1] = -1.0; markw[mm1] = -1.0; markrr[mm1] = 0;
			markx[mm1] = markx[marknum1]; marky[mm1] = marky[marknum1]; markz[mm1] = markz[marknum1];
		}
		mm1++;
	}
	marknum = marknum1;
}

void i3g_movemark(void)
{
	int rc;
	long i;
	static i3g_rock rocks[64];
	glue_grid();
	glue_topography();
	glue_markers();
	/* viscosity, strain and grain state at the old positions (epscalc1 +
	 * viscalc), then the motion */
	int nr = glue_rocks(rocks);
	i3g_mk_params mp;
	glue_mk_params(&mp, 3);
	const float *st12[12] = {exx, eyy, ezz, exy, exz, eyz, sxx, syy, szz, sxy, sxz, syz};
	if ((rc = i3g_markers_update_props(G.mk, rocks, nr, &mp, healrate, st12, pr))) glue_fail("markers_update_props", rc);
	if ((rc = i3g_markers_set_depth_rules(G.mk, zdeep, vdeep, tdeep))) glue_fail("markers_set_depth_rules", rc);
	i3g_move_params mv;
	mv.outgrid = outgrid; mv.markmod = markmod; mv.vyfluid = vyfluid; mv.vymelt = vymelt;
	mv.g[0] = GXKOEF; mv.g[1] = GYKOEF; mv.g[2] = GZKOEF;
	int fluids = 0;
	for (i = 0; i < marknum && !fluids; i++) fluids = markt[i] >= 50 && markt[i] < 100;
	if ((rc = i3g_markers_advect_ex(G.mk, timestep, vx, vy, vz, &mv, fluids ? pr : NULL, fluids ? markd : NULL)))
		glue_fail("markers_advect_ex", rc);
	glue_markers_back();
	glue_immobile();
	/* melt extraction regions cleared, as movemark's end (mark3mg.c:388-392) */
	for (i = 0; i < xnumx * znumz * 5; i++) bufv[i] = 0;
}
