/*
 * i3g — MI355X-native hot path for I3ELVIS-style staggered-grid
 * marker-in-cell geodynamics.  C ABI: plain pointers and sizes only.
 *
 * Conventions (shared with the reference, head3mg.c):
 *   node index n = m3*nx*ny + m1*ny + m2      (y fastest, then x, then z)
 *   y is vertical (depth, positive down); gravity g = (GX, GY, GZ).
 *   unknown index within a level u = n*4 + c, c = 0:P 1:vx 2:vy 3:vz.
 *   level l occupies nodes [mgp[l], mgp[l]+mgn[l]) of every per-node array.
 *   Staggered positions:
 *     vx(m1,m2,m3) at (x[m1],    ymid[m2], zmid[m3])   m2<ny-1, m3<nz-1
 *     vy(m1,m2,m3) at (xmid[m1], y[m2],    zmid[m3])   m1<nx-1, m3<nz-1
 *     vz(m1,m2,m3) at (xmid[m1], ymid[m2], z[m3])      m1<nx-1, m2<ny-1
 *     P (m1,m2,m3) centre of cell [m1-1,m1]x[m2-1,m2]x[m3-1,m3], m1,m2,m3>=1
 *
 * Every function returns I3G_OK (0) or a negative I3G_E* code and never
 * exits the process.  Host pointers are borrowed for the duration of a call.
 * One context per host thread; a context owns one HIP stream.
 */
#ifndef I3G_H
#define I3G_H

#ifdef __cplusplus
extern "C" {
#endif

#define I3G_OK 0
#define I3G_EINVAL -1     /* bad argument / shape */
#define I3G_ENOMEM -2     /* device or host allocation failed */
#define I3G_EDEVICE -3    /* HIP runtime error */
#define I3G_ENOCONV -4    /* solver did not reach the tolerance */
#define I3G_EBC -5        /* unsupported boundary-condition table */
#define I3G_ECAP -6       /* size above a documented capacity */

#define I3G_MAXLEV 10

/* ------------------------------------------------------------------ */
/* Host-side setup (no device needed)                                  */
/* ------------------------------------------------------------------ */

/* Multigrid hierarchy; replaces the reference's gridcheck()
 * (load3mg.c:1162-1283): mgx[l] = (mgx[l-1]-5)/2+5, coarse lines keep the
 * three outer lines of each side at full resolution. */
typedef struct {
	int nlev;                      /* multinum + 1 */
	long mgx[I3G_MAXLEV], mgy[I3G_MAXLEV], mgz[I3G_MAXLEV];
	long mgp[I3G_MAXLEV], mgn[I3G_MAXLEV], mgs[I3G_MAXLEV];
	long ntot;                     /* sum of mgn */
} i3g_hier;

int i3g_hierarchy(long nx, long ny, long nz, int multinum, i3g_hier *h);
/* coarse grid lines of level l (mggx/mggy/mggz in the reference) */
int i3g_level_lines(const i3g_hier *h, int l, const double *gx, const double *gy,
		    const double *gz, double *lx, double *ly, double *lz);

/* One boundary rule of the reference's input deck (in3mg.c:236-407):
 * unknown `var` (0:P 1:vx 2:vy 3:vz 4:T) over the inclusive box
 * [m10,m11]x[m20,m21]x[m30,m31] of level-0 indices obeys
 *   value = cnst + koef * value(shifted by dm1,dm2,dm3)      (koef used
 *   only when a shift is nonzero).  Box ends may be given relative to the
 * last line: set the *_rel flag and the index is (n-1)+offset. */
typedef struct {
	int var;
	long m10, m11, m20, m21, m30, m31;
	int rel10, rel11, rel20, rel21, rel30, rel31;
	double cnst, koef;
	int dm1, dm2, dm3;
} i3g_bc_rule;

/* Free-slip box with fixed top/bottom temperature: the rule set of the
 * reference deck (init.t3c:115-168) with all symmetry coefficients 1.0. */
int i3g_bc_rules_freeslip(double t_top, double t_bottom, i3g_bc_rule *rules,
			  int cap);

/* Expands rules over every level exactly as in3mg.c:315-407.
 * bondm: length 10*nodenum0 (reference layout: 4 per node over all levels,
 * temperature at 8*nodenum0 + node); bondv: [cap][2]; bondn: [cap].
 * Returns the number of records (bondnum) or a negative error. */
long i3g_bc_expand(const i3g_bc_rule *rules, int nrules, const i3g_hier *h,
		   int *bondm, float *bondv, int *bondn, long cap);

/* Seeded synthetic configurations (the benchmark configurations).
 *   which 0: 100 km cube, sphere r=20 km (eta 1e23 / rho 3200) in
 *            eta 1e21 / rho 3300, T linear 273 K -> 1600 K
 *   which 1: 500 km cube, log10 eta = 21 + sum of 32 seeded Gaussian blobs
 *            (sigma 20-60 km, amplitude +-2), rho = 3300 - 50*blob
 *   which 2: 1000 x 500(depth) x 1000 km subduction: lithosphere,
 *            asthenosphere, 45-degree slab, weak interface channel
 * Fills node fields (nu, ro, tk) and stress-point viscosities (nxx at
 * cell centres indexed by the max-corner node, nxy/nxz/nyz at edges) by
 * evaluating the analytic fields at each position. Arrays are level-0 size.
 * params out: [0]=nubeg [1]=nuend. */
int i3g_gen_config(int which, long nx, long ny, long nz, unsigned long long seed,
		   double *gx, double *gy, double *gz, float *nu, float *ro,
		   float *tk, float *nxx, float *nxy, float *nxz, float *nyz,
		   double *params);

/* Markers of the seeded configurations: per_axis^3 per cell at jittered
 * sub-cell centres (seeded), rock type from the layout (configuration 2:
 * 2 asthenosphere, 3 lithosphere, 4 slab, 5 weak channel), temperature from
 * its profile.  With x NULL only the count is returned; else the count or
 * a negative error. */
long i3g_gen_markers(int which, long nx, long ny, long nz, unsigned long long seed, const double *gx,
		     const double *gy, const double *gz, int per_axis, float *x, float *y, float *z, float *tk,
		     signed char *type);

/* ------------------------------------------------------------------ */
/* Run files of the reference (host only)                              */
/* ------------------------------------------------------------------ */

/* The parameter deck: mode.t3c, and the run number in file.t3c; replaces
 * loadconf (load3mg.c:2-180), same token rules (whitespace-separated,
 * tokens starting with '/' are comments) and the same order and units
 * (the output files' time step and timebeg/timeend/convend converted from
 * years to seconds).  The caller provides the output-file table (capacity
 * fl0cap); fl1in becomes the output file of the run number when file.t3c
 * names one (fln3 >= 0).  Errors: I3G_EINVAL (missing file, early end,
 * stp100 <= 0), I3G_ECAP (more output files than fl0cap, multinum or
 * multinumt >= I3G_MAXLEV).  The reference exits instead; it also writes
 * stop.yn, which is the caller's business here. */
typedef struct {
	/* output-file table, caller-owned, fl0cap rows */
	int fl0cap;
	char (*fl0out)[64];
	int *fl0otp;
	int (*fl0cyc)[4];
	double (*fl0stp)[11];
	/* results */
	int fln3;
	char fl1in[64];
	int fl1itp, fl0num;
	int printmod, movemod, tempmod, markmod, gridmod, outgrid, densimod, filesjob;
	double stp100, eroslev, sedilev, waterlev;
	int cyc1max, multinum, multicyc, multinnn[2][I3G_MAXLEV];
	double DIVVMIN, STOKSMIN, DIVVMAX, STOKSMAX;
	double p0koef, p1koef, p2koef, v0koef, v1koef, nubeg, nuend, nukoef;
	int viscmod, spheryn;
	int cyc2max, multinumt, multicyct, multittt[2][I3G_MAXLEV];
	double HEATMIN, t0koef, t1koef, heatdif;
	int frictyn, adiabyn;
	double tkpor, zmpor, vyfluid, vymelt, dmwamin, tdeep, dtdeep, drdeep, zdeep, vdeep, nudeep;
	double dxwater, dywater, dzwater, dserp, yserp, tserp, lambmlt, ydike, dydike, nusseltno;
	double circdepth, circtemp, healrate, meltmin, maxwater, maxpmelt, teclmin, teclmax, xvolc, xtkplut;
	double timebeg, timeend, tempbeg, convend;
} i3g_mode;

int i3g_mode_read(const char *mode_path, const char *file_path, i3g_mode *m);

/* The binary state file (.prn) of loader / saver (load3mg.c:704-1096):
 * byte sizes of int/long/float/double, the header scalars, the rock table
 * (24 columns of rocknum doubles: markn0 markn1 marks0 marks1 marknu markdh
 * markdv markss markmm markll marka0 marka1 markb0 markb1 marke0 marke1
 * markro markbb markaa markcp markkt markkf markkp markht), 16 float fields
 * per node (pr vx vy vz ro nu nxx nxy nxz nyz tk cp et kt ht rr), the grid
 * lines as floats, bondnum BC records (bondm index, record, two floats,
 * partner) and 12 float fields plus the type byte per marker (markx marky
 * markz markk markw markd markv marke markex markft markrr markrn, markt).
 * Arrays are planes: rock[c*rocknum + r], node[c*nodenum + n],
 * mark[c*marknum + i], bond_v[2*q + 0/1].  i3g_prn_read_header reads the
 * scalars (for sizing); i3g_prn_read needs the arrays and a header equal to
 * the file's; i3g_prn_write writes the file byte for byte as saver does.
 * timesum_years is the file's value (the reference converts it to seconds
 * on loading). */
typedef struct {
	long xnumx, ynumy, znumz, mnumx, mnumy, mnumz;
	double xsize, ysize, zsize;
	long pxinit, pyinit, pzinit;
	double pinit, gxkoef, gykoef, gzkoef;
	int rocknum;
	long bondnum, marknum;
	int n0;
	double timesum_years, gridcur, gridtot;
	double *rock;
	float *node;
	float *gx, *gy, *gz;
	long *bond_at, *bond_rec, *bond_partner;
	float *bond_v;
	float *mark;
	signed char *markt;
} i3g_prn;

int i3g_prn_read_header(const char *path, i3g_prn *p);
int i3g_prn_read(const char *path, i3g_prn *p);
int i3g_prn_write(const char *path, const i3g_prn *p);

/* ------------------------------------------------------------------ */
/* Device context and the Stokes solve                                 */
/* ------------------------------------------------------------------ */

typedef struct i3g_ctx i3g_ctx;

typedef struct {
	long nx, ny, nz;               /* xnumx, ynumy, znumz */
	const double *gx, *gy, *gz;    /* level-0 grid lines */
	int multinum;                  /* coarse levels (reference multinum) */
	const int *bondm;              /* reference layout, 10*nodenum0 */
	const float *bondv;            /* [nrec][2] */
	const int *bondn;              /* [nrec] partner+1 (global sol0 index) */
	long nrec;                     /* bondnum + 1 */
} i3g_setup;

/* Creates a context on the current HIP device; replaces the static state
 * of head3mg.c:33-287 for the Stokes/heat/marker path.  The V-cycle's
 * level-0 smoothing is gausmat4's colour phases (at most I3G_SCGS_PHASES,
 * default 6, sweeps) followed by I3G_SCGS_MIX (default 1) sweeps of a
 * coupled cell smoother that has no counterpart in the reference
 * (symmetric coupled Gauss-Seidel: each pressure cell and its six faces
 * corrected together through the cell's Schur complement with the faces'
 * diagonals, cells in two colours, damping I3G_SCGS_W "velocity,pressure",
 * default 0.4,0.5; levels I3G_SCGS_LEVELS "last" or "first,last", default
 * 0).  A stall guard (I3G_SCGS_GUARD V-cycles, default 100, 0 off) switches
 * the coupled sweeps off for the rest of the context's life, and gives the
 * levels their full colour-phase sweeps, when a GCR pass's residual falls
 * to more than I3G_SCGS_GUARD_RATIO (default 0.9) of where it was over
 * that window.  I3G_SCGS=0: colour phases only;
 * I3G_SCGS_MIX=0: coupled sweeps only.  257^3 subduction: 10.7 s (358
 * V-cycles) from a zero start against 16.1 s (366) with the phases only. */
int i3g_create(const i3g_setup *s, i3g_ctx **out);
int i3g_destroy(i3g_ctx *ctx);

/* Viscosity/density inputs of the Stokes operator (the arrays vpiterate
 * reads, move3mg.c:1585). nu/ro at level-0 nodes; nxx at cell centres;
 * nxy/nxz/nyz at edges (index of the edge's own node, as in sxycalc etc.).
 * phys: [0]=GXKOEF [1]=GYKOEF [2]=GZKOEF [3]=nubeg [4]=nuend [5]=viscmod.
 * Coarse-level viscosities are restricted as ronurestrict (move3mg.c:2642).
 * With I3G_COARSE_DIRECT=1 in the environment at i3g_create, this call also
 * assembles the coarsest level's operator and factors it by SVD in FP64
 * (the V-cycle then solves that level directly; grids whose coarsest level
 * has more than 4096 nodes return I3G_ECAP). */
int i3g_set_coefficients(i3g_ctx *ctx, const float *nu, const float *nxx,
			 const float *nxy, const float *nxz, const float *nyz,
			 const float *ro, const double *phys);

/* FP64 residual r = b - A x of the level-0 system, matrix-free, for
 * x = (P, vx, vy, vz) (host arrays, nodenum each).  Replaces evaluating the
 * assembled rows of x/y/zstokserr, continerr and bonderr
 * (move3mg.c:718-1579).  r may be NULL.  norms (may be NULL) receives
 * [0] ||r||_w over non-BC rows, [1] ||b||_w over the same rows, where
 * continuity rows are weighted by eta_cell / mean cell size. */
int i3g_stokes_residual(i3g_ctx *ctx, const double *P, const double *vx,
			const double *vy, const double *vz, double *rP,
			double *rvx, double *rvy, double *rvz, double *norms);

/* i3g_stokes_residual without any allocation, on device buffers: d_x is
 * the level-0 vector (4 planes P, vx, vy, vz of nodenum doubles), d_r the
 * residual out (same layout; NULL: the context's work vector), d_norms 2
 * doubles out ([0] ||r||_w, [1] ||b||_w).  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Its partial sums live in a static
 * device array of 4 MiB per GPU, so the level-0 grid may have at most
 * 262143 blocks of 256 nodes (about 67 million nodes; larger grids return
 * I3G_ECAP).  Enqueued on `stream` (a hipStream_t; NULL = the context's
 * stream) and not synchronised.  Calls that overlap on one GPU race on that
 * array. */
int i3g_stokes_residual_noalloc(i3g_ctx *ctx, const double *d_x, double *d_r,
				double *d_norms, void *stream);

/* Level transfers of the V-cycle, in FP64, on host planes (P, vx, vy, vz;
 * SoA, one plane per unknown kind, level sizes).  `level` is the coarse
 * level (1 <= level < levels), as the reference's argument.
 *
 * i3g_stokes_restrict replaces vprestrict(level) (move3mg.c:3165-3520):
 * coarse right-hand side = weight-normalised trilinear gather of the fine
 * residual per staggered component, fine and coarse BC rows skipped, the P
 * residual scaled by nd_fine/nd_coarse; coarse BC and missing rows get 0.
 * i3g_stokes_prolong replaces vpprolong(level, 0) (move3mg.c:2748-3063):
 * fine += v1koef (velocities) or p1koef (P) times the trilinear gather of
 * the coarse correction, for the existing fine unknowns.
 * Both need i3g_set_coefficients first and overwrite the FP64 multigrid
 * work buffers of the two levels (not the resident solution). */
int i3g_stokes_restrict(i3g_ctx *ctx, int level, const double *const *fine,
			double *const *coarse);
int i3g_stokes_prolong(i3g_ctx *ctx, int level, const double *const *coarse,
		       double *const *fine, double v1koef, double p1koef);

/* The smoother of the V-cycle on host planes of one level (P, vx, vy, vz;
 * SoA), replacing gausmat4's relaxation loop (move3mg.c:3087-3131) with its
 * update rules: velocity rows x += v0koef (b - A x) / a_ii (a_ii the
 * row's own coefficient, the first entry of the reference's row), pressure
 * rows P += p0koef eta_cell (b - A x); boundary rows are not relaxed.
 *
 * i3g_stokes_smooth_phase runs one phase of the colour-ordered sweep:
 * phase 0..5 relaxes velocity component 1 + phase/2 on the nodes with
 * (m1 + m2 + m3) % 2 == phase % 2, phase 6 every pressure row (no row of a
 * phase couples to another row of the same phase, so its result does not
 * depend on the order).  Nothing else changes.
 * i3g_stokes_smooth runs nsweeps whole sweeps (the seven phases in order,
 * then the BC rows x_u = k x_partner, homogeneous as inside the V-cycle)
 * with the product's kernels: variant 0 the context's default, 1 the seven
 * phase launches, 2 one launch per velocity component (the pressure rows in
 * the vz launch; pairs of sweeps out of place, an odd last sweep as
 * phases), 3 the seven phases in z-slabs (phase q of slab [b, b') on planes
 * [b - q, b' - q); slab thickness from I3G_SLAB at create, else 4 planes),
 * 4 the whole FP32 sweep in one launch (k_sweep7: (x, y) tiles marching
 * up through the planes with the phases lagging 0,1,1,2,2,3,3 planes, out
 * of place in pairs; on every level, however few its tiles; FP64 and an
 * odd last sweep as phases), 5 the FP32 sweep in three launches
 * (k_sweep3: one per velocity component, both colours through LDS, P with
 * vz; on every level; pairs out of place).
 * precision 64 or 32 (the planes are rounded to float on the
 * device, and back).  Both need i3g_set_coefficients first and overwrite
 * the level's multigrid work buffers. */
int i3g_stokes_smooth_phase(i3g_ctx *ctx, int level, int phase, int precision,
			    double *const *x, const double *const *b,
			    double v0koef, double p0koef);
int i3g_stokes_smooth(i3g_ctx *ctx, int level, int precision, int nsweeps,
		      int variant, double *const *x, const double *const *b,
		      double v0koef, double p0koef);

typedef struct {
	double rtol;          /* relative weighted residual target (1e-8) */
	int precision;        /* 64: plain FP64 multigrid; 32: FP32 V-cycles
				 inside FP64 defect correction */
	int max_outer;        /* outer (FP64 refinement) iteration cap */
	int krylov;           /* GCR direction storage (directions kept) */
	int max_vcycles;      /* V-cycle cap per outer iteration */
	int nu_pre[I3G_MAXLEV], nu_post[I3G_MAXLEV]; /* sweeps per level */
	int nu_coarse;        /* sweeps on the coarsest level */
	int inner_m;          /* inner FGMRES length per outer direction (0: one V-cycle) */
	double inner_tol;     /* inner FGMRES relative reduction */
	double p0koef, p1koef, v0koef, v1koef;       /* mode.t3c:3032-3036 */
	double inner_rtol;    /* mixed precision: inner reduction per outer */
	int warm_start;       /* 0: start from given fields, 1: zero */
	long pxinit, pyinit, pzinit; double pinit;   /* pressure gauge */
	int verbose;
} i3g_stokes_params;

typedef struct {
	int converged;
	int outer_iters;       /* outer (FP64) iterations */
	int vcycles;           /* V-cycles applied */
	double rel_res;        /* final weighted relative residual */
	double res0;           /* initial relative residual */
	double seconds;        /* device time of the solve */
	double hist[256];      /* relative residual after each outer iteration */
	int nhist;
} i3g_stats;

void i3g_stokes_default_params(i3g_stokes_params *p);

/* Solve to p->rtol from (vx,vy,vz,pr) (used as the initial guess unless
 * warm_start==1) and write the converged fields back; replaces vpiterate
 * (move3mg.c:1585-2636).  Fields are FP64 on the host in this entry point.
 * Pressure is shifted so that P(pxinit,pyinit,pzinit) = pinit as
 * move3mg.c:2248-2269 does. */
int i3g_stokes_solve(i3g_ctx *ctx, double *P, double *vx, double *vy,
		     double *vz, const i3g_stokes_params *p, i3g_stats *st);

/* Device-resident variant used for timing: the solution stays in the
 * context (initial guess: last solution or zero). */
int i3g_stokes_solve_resident(i3g_ctx *ctx, const i3g_stokes_params *p,
			      i3g_stats *st);
int i3g_stokes_get_solution(i3g_ctx *ctx, double *P, double *vx, double *vy,
			    double *vz);
/* The resident solution as the reference's node fields: vx, vy, vz rounded
 * to float and the FP64 pressure (as vpiterate stores them,
 * move3mg.c:2248-2308), into host or device arrays (any may be NULL), and
 * vmax[3] (may be NULL): the largest |v| of each component over its
 * existing unknowns, from which vpiterate limits the time step
 * (move3mg.c:2609-2632). */
int i3g_stokes_fields(i3g_ctx *ctx, float *vx, float *vy, float *vz, double *pr, double *vmax);

/* Sets the resident iterate; unknowns that do not exist on the staggered
 * grid are zeroed and, once coefficients are set, the boundary rows are
 * applied (the guess as the solver starts from it). */
int i3g_stokes_set_solution(i3g_ctx *ctx, const double *P, const double *vx,
			    const double *vy, const double *vz);

/* Residual-optimal start: the resident iterate becomes
 *   x = base + sum_i a_i d_i   (i < k <= 4)
 * with the a_i minimising the weighted FP64 residual (the metric of
 * i3g_stokes_residual), computed on the device from one upload of each
 * vector.  planes: (k+1)*4 host pointers (P, vx, vy, vz of base, then of
 * each d_i; NULL = zero plane).  Missing unknowns are zeroed and boundary
 * rows applied (homogeneous for the d_i).  a_out (k entries) and rel_out
 * (relative residual of the start) may be NULL.  No counterpart in the
 * reference: it serves the learned warm start (the start is never worse
 * than base). */
int i3g_stokes_combine_start(i3g_ctx *ctx, int k, const double *const *planes, double *a_out,
			     double *rel_out);

/* Strain rates and stresses of the resident solution: the value mode of
 * sxxcalc..syzcalc (move3mg.c:3-710) as in the loop at the end of vpiterate
 * (move3mg.c:1691-1764), from the velocities rounded to float (the
 * reference's vx/vy/vz).  Host outputs of nodenum floats, each may be NULL;
 * entries where a component does not exist are 0. */
int i3g_stokes_strain_stress(i3g_ctx *ctx, float *exx, float *eyy, float *ezz, float *exy,
			     float *exz, float *eyz, float *sxx, float *syy, float *szz,
			     float *sxy, float *sxz, float *syz);

/* The reference's own convergence norms of the resident solution
 * (vpiterate, move3mg.c:2405-2596): out[0] stoksum, [1] stoksum1,
 * [2] stoksum2, [3] contsum, [4] contsumb, [5] bondsum, [6] maxvxyz
 * (divvmax > 0 caps it, DIVVMAX).  Momentum rows use the float stresses of
 * i3g_stokes_strain_stress and the double pressure, as the reference. */
int i3g_stokes_native_norms(i3g_ctx *ctx, double divvmax, double *out);

/* Enables (1) or disables (0) HIP-event timing of the named kernels on the
 * context's stream, and clears the collected statistics. */
int i3g_set_timing(i3g_ctx *ctx, int on);

/* Per-kernel timing collected since i3g_set_timing (HIP events on the
 * context's stream): names, average milliseconds per launch, launch counts
 * and algorithmic bytes per launch; returns the count. */
int i3g_kernel_stats(i3g_ctx *ctx, char *names, int name_len, double *avg_ms,
		     long *calls, double *bytes, int cap);

/* ------------------------------------------------------------------ */
/* Heat                                                               */
/* ------------------------------------------------------------------ */

/* Device-resident implicit temperature step on the level-0 grid; replaces
 * the solve part of titerate (heat3mg.c:287-588): heaterr rows (ynerr=1,
 * heat3mg.c:160-225) without adiabatic and shear heating, tbonderr rows
 * from the T block of bondm (heat3mg.c:794-843), relaxation
 * T -= (A T - b)/diag * t0koef in red-black order, convergence when the
 * reference's heat norm (RMS over non-boundary nodes of the row error in
 * K, heat3mg.c:512-560) is <= heatmin. */
typedef struct i3g_heat i3g_heat;

typedef struct {
	double timestep;   /* s; 0 returns at once as the reference */
	double maxtkstep;  /* K: sub-step = maxtkstep / max|H/(ro cp)| (heat3mg.c:358-384) */
	double heatmin;    /* K (HEATMIN) */
	double t0koef;     /* relaxation factor, (0,2) */
	int sweeps;        /* red-black sweeps between norm checks (multittt[0][0]) */
	int max_sweeps;    /* per sub-step; I3G_ENOCONV beyond */
	double timesum, timebeg, timeend, tempend; /* bottom-line plume ramp (heat3mg.c:444-455) */
	int verbose;
} i3g_heat_params;

typedef struct {
	int converged;
	int substeps;
	int sweeps;        /* total over the sub-steps */
	double heatsum;    /* final heat norm (K) */
	double bondsum;    /* final boundary-row norm (K) */
	double seconds;    /* wall time of the call */
	double device_seconds; /* device time of the solve kernels (host copies excluded) */
} i3g_heat_stats;

/* bondm/bondv/bondn of the setup are read for the T block (bondm +
 * 8*nodenum); every surface node must carry a T boundary row. */
int i3g_heat_create(const i3g_setup *s, i3g_heat **out);
int i3g_heat_destroy(i3g_heat *h);
int i3g_heat_default_params(i3g_heat_params *p);
/* Adiabatic and shear heating in heaterr's right-hand side
 * (heat3mg.c:67-160; the reference's adiabyn / frictyn switches) for the
 * following i3g_heat_solve calls: adiabyn 1 adds alpha T rho (g.v), 2 adds
 * alpha T (v.grad P); frictyn 1 adds sigma:eps; both only where
 * rho > 2000 and y < ysize - stp100, with v from vxyzcalc1 at the node and
 * T the temperature at the start of the step.  gk: GXKOEF, GYKOEF, GZKOEF;
 * vx, vy, vz: staggered velocities (float); pr: pressure (double); et: the
 * expansion coefficient (float); sig / eps: 6 float planes each (sxx syy
 * szz sxy sxz syz, exx .. eyz) as the reference stores them.  Fields are
 * copied at the call.  adiabyn = frictyn = 0 switches the terms off. */
int i3g_heat_set_sources(i3g_heat *h, int adiabyn, int frictyn, const double *gk, double ysize,
			 double stp100, const float *vx, const float *vy, const float *vz, const double *pr,
			 const float *et, const float *const *sig, const float *const *eps);
/* tk (in/out, float, level-0 nodes) is advanced by p->timestep; ro, cp,
 * kt, ht are the node fields of ronurecalc.  The start temperature (the
 * reference's tk2) is the caller's input copy. */
int i3g_heat_solve(i3g_heat *h, float *tk, const float *ro, const float *cp, const float *kt,
		   const float *ht, const i3g_heat_params *p, i3g_heat_stats *st);

/* ------------------------------------------------------------------ */
/* Markers                                                            */
/* ------------------------------------------------------------------ */

/* Device-resident marker set on one level-0 grid (the reference's
 * markx/marky/markz/markk/markv/markt arrays, head3mg.c:97-105), kept
 * sorted by cell.  Marker ids give the reference's marker order. */
typedef struct i3g_markers i3g_markers;

/* Rock properties of one marker type: the per-type table of the reference
 * deck (init.t3c ROCKS_DESCRIPTIONS, read by in3mg.c:204-231) as viscalc
 * (mark3mg.c:1114-1681), dencalc1 (:3006-3024) and ronurecalc use it.
 * Density ro*(1-bro*(T-298.15))*(1+aro*(P_bar-1)*1e-3), conductivity
 * (kt + kf/(T+77))*exp(kp*P_bar).  Viscosity: with ss = 0 (or mm = 1)
 * Newtonian, constant (dh = dv = 0) or exp((dh + dv P)/RT); with ss and
 * mm power law (dh > 0: P-T dependent); nu0 < 0 Karato & Wu with ss, ll,
 * s1; dh < 0 the grain-size model (nu0, -dh, dv dislocation; s1, ss, ll
 * diffusion); a0/a1, b0/b1 the brittle strength a + b P at accumulated
 * strains below e0 / above e1.  Zero in the fields after ht leaves only
 * the Newtonian branches. */
typedef struct {
	double nu0, dh, dv;        /* marknu, markdh (J), markdv (J/bar) */
	double n0, n1;             /* markn0, markn1 viscosity limits */
	double ro, bro, aro;       /* markro, markbb, markaa */
	double cp, kt, kf, kp, ht; /* markcp, markkt, markkf, markkp, markht */
	double ss, mm, ll, s1;     /* markss, markmm, markll, marks1 */
	double a0, a1, b0, b1;     /* marka0, marka1, markb0, markb1 */
	double e0, e1;             /* marke0, marke1 */
} i3g_rock;

/* Parameters of the marker -> node transfer (the mode.t3c values that
 * ronurecalc uses) and of movemark's property update: nubeg, nuend,
 * nukoef, densimod (0|1), timestep (> 0: the transfer reuses the stored
 * marker viscosity; the strain step of the property update), vdeep/nudeep
 * viscosity cap below vdeep, viscmod (0, 1, 2).
 * rheology 0: marker types whose ronurecalc branches go beyond viscalc's
 * Newtonian forms (< 2, 7, 8, 16-18, > 20) are refused (I3G_EINVAL), the
 * fields after it are ignored.  rheology 1: the whole viscalc with the
 * strain, grain roughness and dislocation share of the markers; in the
 * transfer the retyping by eroslev / sedilev (to sedtype, the 3 or 4 that
 * ronurecalc derives from sedimnum/sedimcyc) / waterlev, the water/air
 * proportions of markers at the water level (mnumy markers per cell in y),
 * viscalc from the strain rates and stresses of i3g_markers_set_stress,
 * and the phase-transition densities damped between teclmin and teclmax;
 * melt types (21-49) take the properties of the partially molten rock
 * (meltpart0/meltpart11, mark3mg.c:660-686, 1872-2340) from the markers'
 * water content and extracted melt (i3g_markers_set_melt), dry mantle
 * melting stopping above maxpmelt (Pa).  timesum and convend (s) drive the
 * grain growth and the dislocation share.  outgrid: movemark's, for the
 * markers the property update covers (as i3g_move_params.outgrid). */
typedef struct {
	double nubeg, nuend, nukoef;
	int densimod;
	double timestep;
	double vdeep, nudeep;
	int viscmod;
	int rheology;
	double eroslev, sedilev, waterlev;
	int sedtype;
	long mnumy;
	double teclmin, teclmax;
	double timesum, convend;
	int outgrid;
	double maxpmelt;
} i3g_mk_params;

/* movemark's motion parameters (mark3mg.c:61, 146-284): outgrid (0 markers
 * outside the grid move too, 1 they stay, 2 they move only normal to the
 * face they are outside of), markmod (1 one velocity evaluation, else
 * RK4), vyfluid / vymelt the fluid velocities of types 50-99 (vymelt for
 * markd above 1100) and g the gravity (GXKOEF..GZKOEF) that turns them
 * along the pressure gradient. */
typedef struct {
	int outgrid, markmod;
	double vyfluid, vymelt;
	double g[3];
} i3g_move_params;

int i3g_markers_create(long nx, long ny, long nz, const double *gx, const double *gy,
		       const double *gz, long capacity, i3g_markers **out);
int i3g_markers_destroy(i3g_markers *m);
/* id may be NULL (ids 0..n-1); type is the rock type (markt) */
int i3g_markers_upload(i3g_markers *m, long n, const float *x, const float *y, const float *z,
		       const float *tk, const float *eta, const signed char *type, const int *id);
int i3g_markers_download(i3g_markers *m, float *x, float *y, float *z, float *tk, float *eta,
			 signed char *type, int *id);
long i3g_markers_count(const i3g_markers *m);
/* Device time (ms) of the kernels of the last cell_index, sort, to_nodes,
 * update_tk or advect_rk4 call (HIP events on the set's stream; host
 * copies excluded). */
double i3g_markers_last_ms(const i3g_markers *m);

/* Cell of every marker by the reference's binary searches m1serch/m2serch/
 * m3serch (mark3mg.c:1021-1087) after clamping the coordinate to the
 * domain (prcalc1, mark3mg.c:2930-2932); key = m3*nx*ny + m1*ny + m2
 * (mark3mg.c:810-812).  Outputs (host, may be NULL) in resident order. */
int i3g_markers_cell_index(i3g_markers *m, int *m1, int *m2, int *m3, int *key);

/* Stable sort of the resident markers by (cell key, id): counting sort
 * plus an in-cell ordering by id; deterministic.  order (host, may be
 * NULL) receives the previous position of each marker. */
int i3g_markers_sort(i3g_markers *m, long *order);

/* Marker -> node transfer; replaces ronurecalc (mark3mg.c:404-1015)
 * except its melt branch (types 21-49) and the thermodynamic tables
 * (densimod >= 2), which return I3G_EINVAL; prm->rheology selects the
 * Newtonian subset or the whole rheology (see i3g_mk_params).  pr: cell
 * pressures (level-0 node layout, host or device).  Node arrays are
 * in/out (a node with too little marker weight keeps its value, as the
 * reference).  Markers must be sorted; their types, viscosities and
 * strain/grain state are updated as ronurecalc updates them.
 * Deterministic: every node sums its markers in ascending id, the order
 * of the reference's serial loop. */
int i3g_markers_to_nodes(i3g_markers *m, const i3g_rock *rocks, int nrock, const i3g_mk_params *prm,
			 const double *pr, float *nu, float *ro, float *cp, float *kt, float *ht,
			 float *et, float *tk, float *nxx, float *nxy, float *nxz, float *nyz);

/* movemark's depth rules for the following i3g_markers_advect_rk4 calls
 * (mark3mg.c:254-276): a marker below zdeep that sinks becomes type 10 (T
 * raised to tdeep if it already was), any other marker below zdeep type 9;
 * a marker rising through vdeep or zdeep becomes type 10 with T at least
 * tdeep.  Default zdeep = vdeep = 1e30 (off). */
int i3g_markers_set_depth_rules(i3g_markers *m, double zdeep, double vdeep, double tdeep);

/* Accumulated strain of the markers (the reference's marke), in the
 * current marker order (that of i3g_markers_download); upload sets it to 0.
 * Sorting keeps it with its marker. */
int i3g_markers_set_strain(i3g_markers *m, const float *e);
int i3g_markers_get_strain(i3g_markers *m, float *e);

/* epscalc1 (mark3mg.c:2535-2838) at every marker: strain rates, stresses,
 * pressure and dissipation by trilinear interpolation with weights clamped
 * to [0,1] (normal components and P from the cell centres, shear
 * components from their edge nodes).  st12: 12 float node planes exx eyy
 * ezz exy exz eyz sxx syy szz sxy sxz syz as the reference stores them;
 * pr: pressure (double).  out: 14 doubles per marker in the current order
 * (eps xx yy zz xy xz yz, P, sigma xx .. yz, dissipation). */
int i3g_markers_epscalc(i3g_markers *m, const float *const *st12, const double *pr, double *out);

/* movemark's property update ahead of the position update
 * (mark3mg.c:61-145): every marker that moves (prm->outgrid) of type < 50
 * with T > 0 gets eta = viscalc at the strain-rate and stress invariants,
 * dissipation and pressure epscalc1 gives at the marker (with the
 * numerical-diffusion correction of the strain rate, :97-105); a yielding
 * marker's strain grows by timestep * eii, every strain heals by
 * timestep * healrate toward 0.  prm: nubeg, nuend, timestep, rheology
 * (0: types < 2, 7, 8, 16-18, > 20 return I3G_EINVAL; 1: all, with
 * timesum, convend and the grain state), outgrid.  Then
 * i3g_markers_advect_ex moves them. */
int i3g_markers_update_props(i3g_markers *m, const i3g_rock *rocks, int nrock, const i3g_mk_params *prm,
			     double healrate, const float *const *st12, const double *pr);

/* Temperature boundary rows applied at the end of every marker -> node
 * transfer, as ronurecalc calls tkrecalc (mark3mg.c:1007, heat3mg.c:849-899):
 * bondm_t is the level-0 T block of the reference's bondm (bondm +
 * 8*nodenum, nodenum entries, record number or 0), bondv/bondn its record
 * tables (bondv1[r][0..1] at bondv[2r..2r+1], bondn1[r] = partner node + 1
 * or 0), nrec records.  bondm_t NULL switches the update off. */
int i3g_markers_set_tbc(i3g_markers *m, const int *bondm_t, const float *bondv, const int *bondn,
			long nrec);

/* Marker temperature update after the heat solve; replaces the marker
 * part of titerate (heat3mg.c:592-782, interpolation tkcalc1
 * mark3mg.c:2844-2910): markers of type 50..99 inside the grid take the
 * interpolated new temperature; markers of type < 50 strictly inside the
 * grid are initialised from tk2 when markk <= 0 (floor tdeep below zdeep),
 * relaxed towards the grid by the subgrid diffusion (heatdif > 0) whose
 * nodal average is removed again, and given the interpolated change
 * tk - tk2.  tk: temperature after the solve, tk2: before (host, level-0
 * nodes).  Markers must be sorted when heatdif > 0; the nodal averages
 * follow the reference's serial marker order, so the result is
 * deterministic. */
int i3g_markers_update_tk(i3g_markers *m, const float *tk, const float *tk2, const float *kt,
			  const float *ro, const float *cp, double heatdif, double timestep, double zdeep,
			  double tdeep);

/* RK4 advection of the resident markers by dt in the staggered velocity
 * field (vxyzcalc1 + movemark, mark3mg.c:215-284, 2347-2530); markers that
 * leave the grid get tk = eta = 0 (mark3mg.c:280-284).  The motion of
 * i3g_markers_advect_ex with outgrid 0, markmod 0 and no fluid velocity. */
int i3g_markers_advect_rk4(i3g_markers *m, double dt, const float *vx, const float *vy,
			   const float *vz);

/* movemark's motion (mark3mg.c:61, 146-284) with its parameters: which
 * markers move (outgrid), RK4 or one evaluation (markmod), the water
 * velocity of fluid markers (types 50-99: vymelt where markd > 1100, else
 * vyfluid; with vyfluid > 0 and T 0 or above 298 K turned along the
 * pressure gradient of the 8 P nodes around the marker as the reference
 * does it, including its use of the x cell factor for the y and z
 * offsets and its y shift of the two upper nodes; a node index past the
 * last node reads 0), outgrid 2's motion normal to the faces, the depth
 * rules and out-of-grid reset of types < 100, and the motion of immobile
 * markers (>= 100; their reset and the marker regeneration stay with the
 * caller).  pr: P (level-0 nodes, double); markd: the markers' density by
 * marker id (may be NULL when no fluid marker moves or vyfluid and vymelt
 * are 0).  Host or device pointers. */
int i3g_markers_advect_ex(i3g_markers *m, double dt, const float *vx, const float *vy,
			  const float *vz, const i3g_move_params *mp, const double *pr,
			  const float *markd);

/* Grain roughness (markrr) and dislocation share (markrn) of the markers
 * in the current order (as i3g_markers_download); upload sets both to 0,
 * sorting keeps them with their marker.  The grain-size rheology reads and
 * updates them. */
int i3g_markers_set_grain(i3g_markers *m, const float *rr, const float *rn);
int i3g_markers_get_grain(i3g_markers *m, float *rr, float *rn);

/* How the node averages of i3g_markers_to_nodes and i3g_markers_update_tk
 * (the scatters of ronurecalc, mark3mg.c:818-933, and of titerate's
 * subgrid diffusion, heat3mg.c:687-728) sum the markers:
 * 1 (default) cell-segmented: each cell's id-ordered marker segment is
 *   reduced once to FP64 partials of its eight corner nodes, then each node
 *   adds its eight cells' partials in a fixed order; every marker record is
 *   read once and the result does not depend on scheduling;
 * 0 the reference's serial order: each node accumulates the markers of its
 *   eight cells in increasing id, node fields in float as the reference's
 *   arrays (equal to the bit to the reference run on one thread).
 * The two differ by the float accumulation's rounding (about 1e-7 of a
 * field, at most a few 1e-6). */
int i3g_markers_set_order(i3g_markers *m, int order);

/* Water content (markw) and extracted melt fraction (markex) of markers
 * 0..n-1 by marker id (the id given at upload): read by the melt types'
 * properties in the node transfer (meltpart0, mark3mg.c:2152-2340, reads
 * markw[mm1] and markex[mm1]); ids without an entry read 0.  n = 0 clears
 * them.  Host or device pointers. */
int i3g_markers_set_melt(i3g_markers *m, long n, const float *w, const float *ex);

/* Strain rates and stresses (12 float node planes exx eyy ezz exy exz eyz
 * sxx syy szz sxy sxz syz, host or device) that the transfer's viscalc
 * reads (rheology 1, markers without a viscosity or timestep <= 0); NULL
 * clears them (zero strain rates). */
int i3g_markers_set_stress(i3g_markers *m, const float *const *st12);

/* ------------------------------------------------------------------ */
/* Learned warm start                                                  */
/* ------------------------------------------------------------------ */

/* 3D convolutional encoder-decoder (levels of 16/32/64/64 channels, 3x3x3
 * kernels, ReLU, strided convolutions down, node-linear upsampling with
 * additive skips up) mapping two node fields to four.  No counterpart in
 * the reference: it supplies the initial guess that vpiterate takes from
 * the previous step's vx/vy/vz/pr (move3mg.c:1585-1700).  BF16 matrix-core
 * products with FP32 accumulation; FP32 parameters and Adam.
 * Grids: nz, nx, ny = 8k+1 (three halvings).  Fields are node-major with
 * channels fastest: in [nodes][2], out/target [nodes][4]; the caller
 * chooses their scaling.  Parameter vector: per layer W [27][cout][cin]
 * (tap 9 tz + 3 tx + ty) then b [cout], layers in forward order. */
typedef struct i3g_net i3g_net;

int i3g_warmstart_create(long nz, long nx, long ny, unsigned long long seed, i3g_net **out);
int i3g_warmstart_destroy(i3g_net *net);
long i3g_warmstart_nparams(const i3g_net *net);
int i3g_warmstart_get_params(i3g_net *net, float *params);
int i3g_warmstart_set_params(i3g_net *net, const float *params);
int i3g_warmstart_infer(i3g_net *net, const float *in, float *out);
/* The same on device buffers (n = nodes of the network's grid, else
 * I3G_EINVAL): d_in n x 2 features, d_out n x 4 outputs, node order.  No
 * reference counterpart: it keeps the learned start of the Stokes solve
 * on the device (i3g_stokes_learned_start). */
int i3g_warmstart_infer_device(i3g_net *net, long n, const float *d_in, float *d_out);

/* Learned start of the resident Stokes iterate, all on the device: network
 * features from the context's level-0 viscosity and density, inference,
 * the lithostatic pressure (rho g_y integrated down each column at the
 * pressure points, zero at gauge = {m1, m2, m3}), then the residual-optimal
 * combination x = lithostatic + a0 (network P * pscale) + a1 (network
 * velocities * vscale), as i3g_stokes_combine_start; the zero start is
 * taken instead when its residual is smaller.  The network must be on the
 * context's level-0 grid.  info[6]: a0, a1, relative residual of the
 * combination, of the lithostatic start alone, of the zero start, and 1 if
 * the combination was taken (0: zero start).  No reference counterpart. */
int i3g_stokes_learned_start(i3g_ctx *ctx, i3g_net *net, double pscale, double vscale, const long gauge[3],
			     double *info);
/* The same with the fallback chosen: 0 the zero start, 1 the resident
 * iterate as it was before the call (in a time loop the previous step's
 * solution, vpiterate's own start, move3mg.c:1973-2019); info[4] is then
 * the resident iterate's residual and info[5] = 0 when it was kept. */
int i3g_stokes_learned_start_ex(i3g_ctx *ctx, i3g_net *net, double pscale, double vscale,
				const long gauge[3], int fallback, double *info);
/* gradient of the mean squared error of one sample (grad may be NULL) */
int i3g_warmstart_grad(i3g_net *net, const float *in, const float *target, double *loss, float *grad);
/* one Adam step (beta 0.9/0.999, eps 1e-8) on the mean gradient of a batch */
int i3g_warmstart_train_step(i3g_net *net, int batch, const float *in, const float *target, double lr,
			     double *loss);

/* ------------------------------------------------------------------ */
/* The time-step loop                                                  */
/* ------------------------------------------------------------------ */

/* The per-step sequence of the reference's main loop (i3mg.c:80-151) on one
 * GPU with every field resident on the device: the Stokes solve
 * (vpiterate: the marker time step limited to maxxystep / |v|max), the heat
 * step and the marker temperature update (titerate), RK4 advection
 * (movemark), and the marker -> node transfer after a cell re-sort
 * (ronurecalc with tkrecalc).  Node fields move between the stages as
 * device arrays; the host sees only scalars.  No reference counterpart as
 * one call: it is the caller i3mg.c restated over the C ABI above. */
typedef struct i3g_model i3g_model;

typedef struct {
	i3g_mk_params mk;         /* ronurecalc (its timestep is ignored: the
				   * marker viscosity is recomputed every step) */
	i3g_stokes_params stokes; /* vpiterate's solve; the start is the
				   * previous step's solution */
	i3g_heat_params heat;     /* titerate; timestep set per step */
	double maxtmstep;         /* the step before the displacement limit */
	double maxxystep;         /* largest marker displacement per step */
	double heatdif;           /* marker subgrid diffusion coefficient */
	double pscale, vscale;    /* learned start output scales */
	int learned;              /* 1: learned start (network given) with the
				   * previous solution as its fallback */
} i3g_model_params;

typedef struct {
	double seconds;           /* wall time of the step */
	double t_stokes, t_heat, t_mktk, t_rk4, t_sort, t_transfer; /* per stage */
	int vcycles, outer;
	double rel_res, timestep, timesum;
	int learned_taken;        /* 1: the learned start was used */
	int heat_substeps, heat_sweeps;
} i3g_model_stats;

/* grid, BC tables (velocity and temperature) and marker capacity */
int i3g_model_create(const i3g_setup *s, long marker_capacity, i3g_model **out);
int i3g_model_destroy(i3g_model *m);
/* markers (host arrays; ids 0..n-1 in the given order) and the rock table */
int i3g_model_set_markers(i3g_model *m, long n, const float *x, const float *y, const float *z,
			  const float *tk, const signed char *type, const i3g_rock *rocks, int nrock);
/* the physics of the Stokes operator: gravity and viscosity limits,
 * phys as i3g_set_coefficients */
int i3g_model_set_physics(i3g_model *m, const double *phys);
/* initial marker -> node transfer (needs the pressure for densities; NULL
 * = zero), as the node fields a reference run starts from */
int i3g_model_init(i3g_model *m, const i3g_model_params *p, const double *pr);
int i3g_model_step(i3g_model *m, const i3g_model_params *p, i3g_net *net, i3g_model_stats *st);
/* node fields (host or device arrays, any may be NULL): velocities and
 * pressure of the last solve, temperature and viscosity of the last
 * transfer */
int i3g_model_get(i3g_model *m, float *vx, float *vy, float *vz, double *pr, float *tk, float *nu);
/* the model's marker set (for downloads) and Stokes context */
i3g_markers *i3g_model_markers(i3g_model *m);
i3g_ctx *i3g_model_stokes(i3g_model *m);

/* ------------------------------------------------------------------ */
/* Initial condition generator and ParaView output                    */
/* ------------------------------------------------------------------ */

/* A rock body of the reference deck (init.t3c rock boxes, read by
 * in3mg.c:550-575).  shape 0: a hexahedron whose faces are bilinear
 * surfaces x = a0 + a1 y + a2 z + a3 y z (left: corners 0 1 4 5, right:
 * 2 3 6 7), y = .. (top 0 2 4 6, bottom 1 3 5 7), z = .. (front 0-3, back
 * 4-7); shape 1: a spherical shell around corner 0, the distance scaled by
 * corner 7's x, y, z, between corner 2's x and corner 5's x; shape 2: the
 * same for a cylinder along y (its half length scaled by corner 7's y).
 * Coordinates are fractions of the domain size. */
typedef struct {
	int shape;
	int type;       /* rock type given to the markers inside (>= 100: immobile) */
	double ex, e;   /* melt extraction (markex) and strain (marke) given to them */
	double x[8], y[8], z[8];
} i3g_rock_box;

/* Paint the resident markers with the rock bodies in order, as in3mg.c:
 * 548-739 does (a later body wins): type and strain on the device, and for
 * hexahedra the temperature slot markk, set to the marker's x (immobile
 * types >= 100, whose origin it holds) or 0.  ex, od, ow: the markers'
 * markex, markd, markw in the current order (in/out, host or device, any
 * may be NULL); hexahedra set markd/markw to the immobile origin (y, z) or
 * 0.  The face coefficients come from the reference's 4x4 elimination
 * without pivoting (gausmat, gaus3mg.c:23-62), restated, so that markers on
 * a face fall on the same side. */
int i3g_markers_paint(i3g_markers *m, double xsize, double ysize, double zsize,
		      const i3g_rock_box *boxes, int nbox, float *ex, float *od, float *ow);

/* A temperature box of the reference deck (init.t3c, read by in3mg.c:
 * 806-829): type 0 trilinear between the corner temperatures t0..t7; 1/4
 * half-space cooling T = t1 - erfc(d / 2 sqrt(kappa age)) (t1 - t0) + t7 d
 * with the age bilinear from t2..t5 and kappa t6; 5/6 a transition
 * bilinear between the field's own columns at the box corners (x0, z0),
 * (x2, z0), (x5, z7), (x7, z7); 2 a spherical shell around (x0, y0, z0),
 * distances scaled by (x7, y7, z7), between x2/xsize and x5/xsize, T from
 * t0 to t1; 3 the cylinder along y.  The corner coordinates in metres. */
typedef struct {
	int type;
	double x0, x2, y0, y1, y2, y3, z0, x5, x7, y4, y5, y6, y7, z7;
	double t[8];
} i3g_temp_box;

/* Node temperatures from the boxes in order (in3mg.c:830-951), on the
 * device: tk (level-0 nodes, host or device, in/out).  xsize etc. are the
 * deck's domain sizes (types 2, 3 scale by them).  The transition types
 * read the field as the reference's serial loop leaves it: one thread per
 * y line walks its nodes in the reference's order. */
int i3g_temperature_boxes(long nx, long ny, long nz, const double *gx, const double *gy,
			  const double *gz, double xsize, double ysize, double zsize,
			  const i3g_temp_box *boxes, int nbox, float *tk);

/* The rock and temperature boxes of a deck in the reference's format
 * (init.t3c): the sections after its second and third '~' (ffscanf1's
 * tokens: whitespace separated, tokens starting with '/' skipped; 'm'
 * prefixes metres, other coordinates are fractions of the sizes; 's'/'c'
 * type prefixes select the shell shapes).  At most maxrb / maxtb boxes. */
int i3g_deck_boxes(const char *path, double xsize, double ysize, double zsize,
		   i3g_rock_box *rb, int maxrb, int *nrb, i3g_temp_box *tb, int maxtb, int *ntb);

/* ParaView output as the reference writes it (i3vis_output_paraview.c,
 * binary appended, little endian): i3g_write_vtr_nodal writes the
 * RectilinearGrid of the node fields (x fastest in the file; host arrays
 * in the node order m3, m1, m2; f64[q] != 0 marks a double field, f64 may
 * be NULL), as saver_paraview writes "viscosity", "temperature",
 * "density", "vy", "vx", "vz" (float) and "pressure" (double);
 * i3g_write_vtr_composition the cell
 * composition map of the markers on an mx x my x mz uniform grid
 * (i3vis_generate_compo_map, including its claim rule and neighbour fill,
 * :1111-1314); i3g_pvd_open / i3g_pvd_append the time collection. */
int i3g_write_vtr_nodal(const char *path, long nx, long ny, long nz, const double *gx,
			const double *gy, const double *gz, int nfield, const char *const *names,
			const void *const *fields, const int *f64);
int i3g_write_vtr_composition(const char *path, double xsize, double ysize, double zsize, long mx,
			      long my, long mz, long n, const float *x, const float *y,
			      const float *z, const signed char *type);
int i3g_pvd_open(const char *path);
int i3g_pvd_append(const char *path, double time, const char *datafile, const char *dir);

const char *i3g_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif
