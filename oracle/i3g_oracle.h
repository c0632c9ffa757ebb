/*
 * i3g oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the reference algorithm
 * for the Stokes hot path.  Only tests/ and __graft_entry__.smoke() (to
 * check results) and bench.py's CPU-baseline leg (the timed CPU side) may
 * load it.
 * Vectors use the reference's sol0 layout: unknown (node n, comp c) at
 * n*4+c with c = 0:P 1:vx 2:vy 3:vz; a level's node n is level-local.
 * Boundary tables are the reference's global layout (bondm over all levels,
 * bondn partner+1 as a global sol0 index); `mgp` is the level's node offset.
 */
#ifndef I3G_ORACLE_H
#define I3G_ORACLE_H

typedef struct {
	long nx, ny, nz;   /* level dims */
	long mgp;          /* node offset of the level in the global tables */
	int level;         /* 0 = finest */
	const double *gx, *gy, *gz; /* level lines */
} orl_t;

/* Effective viscosity at the stress points (move3mg.c:36-48 sxx/syy/szz
 * and continerr, :402-414 sxy, :517-529 sxz, :632-644 syz). nxx..nyz are
 * used only on level 0 (may be NULL otherwise). Output arrays have N
 * entries; positions where the stress does not exist are set to 0. */
void or_eta(const orl_t *L, const float *nu, const float *nxx, const float *nxy,
	    const float *nxz, const float *nyz, int viscmod, double nubeg,
	    double nuend, double *etaC, double *etaXY, double *etaXZ,
	    double *etaYZ);

/* Momentum right-hand side on level 0 from node densities (move3mg.c:751
 * and the y/z analogues); continuity and BC entries get 0/const. Output
 * 4N interleaved; BC rows get their constant (bonderr ui1[0]). */
void or_rhs(const orl_t *L, const float *ro, const double *g, const int *bondm,
	    const float *bondv, double *b);

/* r = b - A x for every existing unknown of the level (rows of
 * x/y/zstokserr, continerr and bonderr with ynerr=0, move3mg.c:718-1579);
 * missing unknowns get r = 0. */
void or_residual(const orl_t *L, const double *etaC, const double *etaXY,
		 const double *etaXZ, const double *etaYZ, const int *bondm,
		 const float *bondv, const int *bondn, const double *b,
		 const double *x, double *r);

/* Weighted norms: out[0] = ||r||, out[1] = ||b|| over non-BC rows with
 * continuity rows scaled by etaC / mean cell size. */
void or_wnorms(const orl_t *L, const double *etaC, const int *bondm,
	       const double *b, const double *r, double *out);

/* Value-mode strain rates and stresses (move3mg.c:1691-1764 loop) from FP32
 * fields; eps/sig arrays are N each (xx,yy,zz,xy,xz,yz). */
void or_stress(const orl_t *L, const double *etaC, const double *etaXY,
	       const double *etaXZ, const double *etaYZ, const float *vx,
	       const float *vy, const float *vz, float *eps6, float *sig6);

/* Coarse viscosity restriction, scatter form (move3mg.c:2642-2741). */
void or_ronurestrict(const orl_t *F, const orl_t *C, const float *nu_f,
		     float *nu_c, int viscmod);

/* Residual restriction, scatter form (move3mg.c:3165-3520).  r_f: fine 4Nf
 * interleaved; nd_f/nd_c: cell viscosities (float, as stored by continerr);
 * b_c: coarse 4Nc interleaved output. */
void or_vprestrict(const orl_t *F, const orl_t *C, const int *bondm,
		   const double *r_f, const float *nd_f, const float *nd_c,
		   double *b_c);

/* Correction prolongation (move3mg.c:2748-3063): x_f += P x_c. */
void or_vpprolong(const orl_t *F, const orl_t *C, const double *x_c,
		  double *x_f, double v1koef, double p1koef);

/* One reference-style Gauss-Seidel relaxation sweep (gausmat4 update rules,
 * move3mg.c:3087-3131) in lexicographic node order, matrix-free. */
void or_gs_sweep(const orl_t *L, const double *etaC, const double *etaXY,
		 const double *etaXZ, const double *etaYZ, const int *bondm,
		 const float *bondv, const int *bondn, const double *b, double *x,
		 double p0koef, double v0koef);

/* One phase of the colour-ordered sweep with the gausmat4 update rules
 * (move3mg.c:3100-3128): phase 0..5 = velocity component 1 + phase/2 on the
 * nodes with (m1+m2+m3) % 2 == phase % 2, phase 6 = every pressure row.
 * BC rows untouched; x updated in place from A x of the state before. */
void or_colour_phase(const orl_t *L, const double *etaC, const double *etaXY,
		     const double *etaXZ, const double *etaYZ, const int *bondm,
		     const double *b, double *x, int phase, double p0koef, double v0koef);

/* ---- markers (test infrastructure) ---------------------------------- */

typedef struct {
	double nu0, dh, dv, n0, n1, ro, bro, aro, cp, kt, kf, kp, ht;
} or_rock_t;

/* m1serch/m2serch/m3serch after clamping to the domain (mark3mg.c:1021-1087,
 * 2930-2936); out[3*i..] = (m1, m2, m3) */
void or_mk_cells(long n, const float *x, const float *y, const float *z, const double *gx, long nx,
		 const double *gy, long ny, const double *gz, long nz, int *out);

/* ronurecalc (mark3mg.c:404-1015) for Newtonian rock types without air/
 * water, eclogite or melt branches: serial marker loop, float accumulators,
 * double weights, normalisation.  fld: 11 planes of N (nu ro cp kt ht et tk
 * nxx nxy nxz nyz), in/out.  eta (markv) in/out. */
void or_ronurecalc(long n, const float *x, const float *y, const float *z, const float *tk, float *eta,
		   const signed char *type, const double *gx, long nx, const double *gy, long ny, const double *gz,
		   long nz, const double *pr, const or_rock_t *rocks, double nubeg, double nuend, double nukoef,
		   int densimod, double timestep, double vdeep, double nudeep, float *fld);

/* vxyzcalc1 (mark3mg.c:2347-2530): v[3] at a point */
void or_vxyz(const float *vx, const float *vy, const float *vz, const double *gx, long nx, const double *gy,
	     long ny, const double *gz, long nz, double x, double y, double z, double *v);

/* movemark RK4 position update (mark3mg.c:215-284) for markers of type <
 * 100 that move, with the depth rules (zdeep, vdeep, tdeep; 1e30 = off,
 * mark3mg.c:254-276); out-of-grid markers get tk = eta = 0. */
void or_rk4(long n, float *x, float *y, float *z, float *tk, float *eta, signed char *type,
	    const float *vx, const float *vy, const float *vz, const double *gx, long nx, const double *gy,
	    long ny, const double *gz, long nz, double dt, double zdeep, double vdeep, double tdeep);

/* epscalc1 (mark3mg.c:2535-2838): strain rates, stresses, pressure and
 * dissipation at a point by trilinear interpolation with weights clamped
 * to [0,1]: normal components and P from the cell centres, the shear
 * components from their edge nodes.  st12: exx eyy ezz exy exz eyz, sxx
 * .. syz planes.  out[0..5] eps xx yy zz xy xz yz, out[6] P,
 * out[7..12] sigma, out[13] dissipation sigma:eps. */
void or_epscalc(const float *const *st12, const double *pr, const double *gx, long nx, const double *gy, long ny,
		const double *gz, long nz, double x, double y, double z, double *out);

/* movemark's marker properties ahead of the position update
 * (mark3mg.c:61-145) for the Newtonian rock types: for the markers that
 * move, epscalc1 at the marker gives the strain-rate invariant and the
 * pressure; markers of type < 50 with T > 0 get eta = viscalc(T, P),
 * viscalc marks them as not yielding (e = -|e|, no brittle strength for
 * these rocks), so e is not grown by dt*eii and heals by dt*healrate
 * toward zero. */
void or_mk_props(long n, const float *x, const float *y, const float *z, const signed char *type, const float *tk,
		 float *eta, float *e, const or_rock_t *rocks, const float *const *st12, const double *pr,
		 const double *gx, long nx, const double *gy, long ny, const double *gz, long nz, double nubeg,
		 double nuend, double dt, double healrate);

/* ---- full rheology (test infrastructure) ------------------------------- */

/* every per-rock value viscalc and ronurecalc read (in3mg.c:204-231):
 * the thirteen of or_rock_t, then markss, markmm, markll, marks1, marka0,
 * marka1, markb0, markb1, marke0, marke1 */
typedef struct {
	double nu0, dh, dv, n0, n1, ro, bro, aro, cp, kt, kf, kp, ht;
	double ss, mm, ll, s1, a0, a1, b0, b1, e0, e1;
} or_rockx_t;

/* viscalc's globals: nubeg, nuend, timestep, timesum, convend (s) */
typedef struct {
	double nubeg, nuend, timestep, timesum, convend;
} or_visprm_t;

/* ronurecalc's globals */
typedef struct {
	double nukoef;
	int densimod;
	double vdeep, nudeep;
	double eroslev, sedilev, waterlev;
	int sedtype; /* the sediment type of this call (3 or 4, from sedimnum/sedimcyc) */
	long mnumy;  /* markers per cell along y */
	double teclmin, teclmax;
	double maxpmelt; /* Pa: no dry mantle melting above it */
} or_mkx_t;

/* movemark's globals of the motion */
typedef struct {
	int outgrid, markmod;
	double vyfluid, vymelt;
	double g[3]; /* GXKOEF, GYKOEF, GZKOEF */
} or_movex_t;

double or_viscalc(const or_rockx_t *R, int mm2, double mtk, double mpb, double eii, double sii, double rii,
		  double dis, float *me, float *mrr, float *mrn, const or_visprm_t *P);
void or_mk_props_x(long n, const float *x, const float *y, const float *z, const signed char *type, const float *tk,
		   float *eta, float *e, float *rr, float *rn, const or_rockx_t *rocks, const float *const *st12,
		   const double *pr, const double *gx, long nx, const double *gy, long ny, const double *gz, long nz,
		   const or_visprm_t *P, double healrate, int outgrid);
int or_ronurecalc_x(long n, const float *x, const float *y, const float *z, const float *tk, float *eta,
		    signed char *type, float *e, float *rr, float *rn, const double *gx, long nx, const double *gy,
		    long ny, const double *gz, long nz, const double *pr, const float *const *st12,
		    const or_rockx_t *rocks, const or_visprm_t *P, const or_mkx_t *M, const float *mw,
		    const float *mex, float *fld);
/* node sums of or_ronurecalc / or_ronurecalc_x: 0 float as the reference
 * (default), 1 double */
void or_set_sum64(int on);
/* meltpart11 (mark3mg.c:1872-2070): melt fraction and latent heat (J/kg)
 * of rock type mm2 with water content mwa at mtk (K), mpb (bar) */
void or_meltpart11(int mm2, double mtk, double mpb, double mwa, double maxpmelt, double *xmelt, double *hlat);
/* meltpart0 (mark3mg.c:2152-2340): eps1[20..26] of a marker of melt type
 * mm2 (> 20) with water content mwa and extracted melt mex; returns 1 for
 * mm2 > 21, 0 (and zeros) otherwise */
int or_meltpart0(const or_rockx_t *rocks, int mm2, double mtk, double mpb, double mwa, double mex, int densimod,
		 double teclmin, double teclmax, double maxpmelt, double *e20_26);
void or_rk4_x(long n, float *x, float *y, float *z, float *tk, float *eta, signed char *type, const float *md,
	      const float *vx, const float *vy, const float *vz, const double *pr, const double *gx, long nx,
	      const double *gy, long ny, const double *gz, long nz, double dt, double zdeep, double vdeep,
	      double tdeep, const or_movex_t *M);

/* tkrecalc (heat3mg.c:846-899): temperature boundary rows on level 0 from
 * a copy of tk (reference bondm layout, T records at 8*nodenum). */
void or_tkrecalc(long nodenum, const int *bondm, const float *bondv, const int *bondn, float *tk);

/* ---- heat (test infrastructure) -------------------------------------- */

typedef struct {
	long nx, ny, nz;
	const double *gx, *gy, *gz;
	const float *ro, *cp, *kt, *ht; /* node fields */
	const float *tk0;               /* temperature at the start of the sub-step */
	double dt;                      /* sub-step */
	const int *bondm_t;             /* level-0 T block of bondm (nodenum entries) */
	const float *bondv;             /* bondv1[r][0..1] at [2r..2r+1] */
	const int *bondn;               /* bondn1[r] = partner node + 1 or 0 */
	double timesum, timebeg, timeend, tempend; /* bottom-line plume override */
	const float *src;               /* heaterr source as stored in tk1 (float), NULL = ht */
} or_heat_t;

/* heaterr right-hand-side mode (ynerr=0, heat3mg.c:54-160) at every
 * non-boundary level-0 node: src = radiogenic heat + adiabatic heating
 * (adiabyn 1: alpha T rho (g.v), 2: alpha T (v.grad P)) + shear heating
 * (frictyn: sigma:eps), the two terms only where rho > 2000 and
 * y < ysize - stp100; rate = src / (ro cp).  tk0 is the temperature at the
 * start of the step, v at the node is vxyzcalc1's, s6/e6 are sxx syy szz
 * sxy sxz syz and exx .. eyz.  src is stored as float, as tk1; rate is the
 * double quotient heaterr returns.  Boundary nodes get 0. */
void or_heat_source(long nx, long ny, long nz, const double *gx, const double *gy, const double *gz,
		    const float *ro, const float *cp, const float *ht, const float *tk0, const float *et,
		    const float *vx, const float *vy, const float *vz, const double *pr, const float *const *s6,
		    const float *const *e6, const int *bondm_t, int adiabyn, int frictyn, const double *gk,
		    double ysize, double stp100, float *src, double *rate);

/* Rows of the level-0 heat system (7 slots per node: cnt entries of
 * cols/vals at 7*node) and right-hand side, heat3mg.c:160-225, 818-841. */
void or_heat_rows(const or_heat_t *H, int *cnt, int *cols, double *vals, double *rhs);
/* max |H/(ro cp)| over non-boundary nodes (heat3mg.c:327-368) */
double or_heat_maxrate(const or_heat_t *H);
/* out[0] heatsum, out[1] bondsum of titerate (heat3mg.c:512-560) */
void or_heat_norms(const or_heat_t *H, const double *x, double *out);
/* marker temperature update after the heat solve (heat3mg.c:592-782) */
void or_mk_tk_update(long n, const float *x, const float *y, const float *z, const signed char *type, float *markk,
		     const double *gx, long nx, const double *gy, long ny, const double *gz, long nz, const float *tk,
		     const float *tk2, const float *kt, const float *ro, const float *cp, double heatdif, double dt,
		     double zdeep, double tdeep);

/* ---- native norms (test infrastructure) -------------------------------- */

/* Residual-mode rows of vpiterate's norm loop (move3mg.c:2405-2560):
 * rows[k][u][0..1] for kind k = 0 continerr (u = 4n), 1..3 x/y/zstokserr
 * (return value, eps1[0] = value * skf1) on non-BC rows, 4 bonderr of the
 * BC velocity rows; NaN elsewhere.  st12: exx..eyz, sxx..syz float planes;
 * sol: interleaved double solution (for bonderr). */
void or_err_rows(const orl_t *L, const float *nu, const float *ro, const double *g, double nubeg, double nuend,
		 const float *st12, const double *pr, const float *vx, const float *vy, const float *vz,
		 const int *bondm, const float *bondv, const int *bondn, const double *sol, double *rows);

#endif
