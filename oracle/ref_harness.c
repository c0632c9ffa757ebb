/*
 * Reference harness (test infrastructure only; never part of the product).
 *
 * The recipe in oracle/Makefile feeds this file to gcc together with the
 * reference's own sources, read where they lie under the reference checkout:
 * the reference's global-state header is streamed through sed with smaller
 * static capacities (its stock capacities need ~78 GB of BSS) and this file
 * then #includes the reference's solver sources (load3mg.c, move3mg.c,
 * mark3mg.c, gaus3mg.c, heat3mg.c).  No reference source is copied into the
 * repository; the only output is oracle/_ref/libi3ref.so.
 *
 * The entry points below only set up the reference's globals and call the
 * reference's own functions.  The loops that walk the unknowns in
 * refh_assemble_level() mirror the assembly loop of vpiterate
 * (move3mg.c:2031-2157): which unknown exists, and whether it is a
 * boundary-condition row or a stencil row.
 */

#include "load3mg.c"
#include "move3mg.c"
#include "mark3mg.c"
#include "gaus3mg.c"
#include "heat3mg.c"

/* Static capacities of this build (returned so that ctypes can view arrays). */
void refh_caps(long *out)
{
	out[0] = MAXXLN; out[1] = MAXYLN; out[2] = MAXZLN;
	out[3] = MAXNOD; out[4] = MAXPPP; out[5] = MAXPAR;
	out[6] = MAXBON; out[7] = MAXMRK; out[8] = MAXNNN;
	out[9] = MAXMAT; out[10] = MAXTMR;
}

/* Grid + multigrid hierarchy through the reference's gridcheck()
 * (load3mg.c:1162-1283). */
int refh_setup_grid(long nx, long ny, long nz, const double *x, const double *y,
		    const double *z, int nlev)
{
	long i;
	if (nx > MAXXLN || ny > MAXYLN || nz > MAXZLN) return -1;
	if (nx * ny * nz * 2 > MAXNOD) return -2;
	xnumx = nx; ynumy = ny; znumz = nz;
	for (i = 0; i < nx; i++) gx[i] = x[i];
	for (i = 0; i < ny; i++) gy[i] = y[i];
	for (i = 0; i < nz; i++) gz[i] = z[i];
	xsize = gx[nx - 1]; ysize = gy[ny - 1]; zsize = gz[nz - 1];
	multinum = multimax = nlev;
	marknum = 0; rocknum = 0;
	printmod = 0;
	gridcheck();
	return 0;
}

/* One row of the assembled system for unknown (mgi,m1,m2,m3,comp) exactly as
 * vpiterate stores it.  Returns the entry count, or -1 when that unknown does
 * not exist on the staggered grid.  Column indices are global sol0 indices. */
static long refh_row(int mgi, long m1, long m2, long m3, int comp, long *un1,
		     double *ui1, double *eps1)
{
	long mc = (mgp[mgi] + m3 * mgxy[mgi] + m1 * mgy[mgi] + m2) * 4;
	switch (comp) {
	case 0:
		if (!(m1 && m2 && m3)) return -1;
		if (!bondm[mc + 0]) continerr(m1, m2, m3, 0, mgi, eps1, un1, ui1);
		else bonderr(mc + 0, 0, eps1, un1, ui1);
		break;
	case 1:
		if (!(m2 < mgy[mgi] - 1 && m3 < mgz[mgi] - 1)) return -1;
		if (!bondm[mc + 1]) xstokserr(m1, m2, m3, 0, mgi, eps1, un1, ui1);
		else bonderr(mc + 1, 0, eps1, un1, ui1);
		break;
	case 2:
		if (!(m1 < mgx[mgi] - 1 && m3 < mgz[mgi] - 1)) return -1;
		if (!bondm[mc + 2]) ystokserr(m1, m2, m3, 0, mgi, eps1, un1, ui1);
		else bonderr(mc + 2, 0, eps1, un1, ui1);
		break;
	default:
		if (!(m1 < mgx[mgi] - 1 && m2 < mgy[mgi] - 1)) return -1;
		if (!bondm[mc + 3]) zstokserr(m1, m2, m3, 0, mgi, eps1, un1, ui1);
		else bonderr(mc + 3, 0, eps1, un1, ui1);
		break;
	}
	return un1[0];
}

/* Assemble one level into CSR (rows = 4*mgn[mgi] local unknowns in sol0
 * order).  rowptr has 4*mgn+1 entries; cols/vals capacity 17 per row.
 * Missing unknowns get an empty row and rhs 0. */
long refh_assemble_level(int mgi, long *rowptr, long *cols, double *vals,
			 double *rhs)
{
	long m1, m2, m3, k, nnz = 0, r;
	int c;
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	rowptr[0] = 0;
	for (m3 = 0; m3 < mgz[mgi]; m3++)
	for (m1 = 0; m1 < mgx[mgi]; m1++)
	for (m2 = 0; m2 < mgy[mgi]; m2++)
	for (c = 0; c < 4; c++) {
		r = (m3 * mgxy[mgi] + m1 * mgy[mgi] + m2) * 4 + c;
		un1[0] = 0; ui1[0] = 0;
		long cnt = refh_row(mgi, m1, m2, m3, c, un1, ui1, eps1);
		if (cnt > 0) {
			for (k = 1; k <= cnt; k++) {
				cols[nnz] = un1[k];
				vals[nnz] = ui1[k];
				nnz++;
			}
			rhs[r] = ui1[0];
		} else {
			rhs[r] = 0;
		}
		rowptr[r + 1] = nnz;
	}
	return nnz;
}

/* Assemble every level into the reference's own row store through gausmat2,
 * as vpiterate does before its V-cycles (move3mg.c:2029-2159). */
void refh_store_all_levels(void)
{
	long m1, m2, m3, mcmax1;
	int mgi, c;
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	for (mgi = 0; mgi <= multimax; mgi++)
	for (m3 = 0; m3 < mgz[mgi]; m3++)
	for (m1 = 0; m1 < mgx[mgi]; m1++)
	for (m2 = 0; m2 < mgy[mgi]; m2++) {
		mcmax1 = mgp[mgi] + m3 * mgxy[mgi] + m1 * mgy[mgi] + m2;
		for (c = 0; c < 4; c++) {
			un1[0] = 0; ui1[0] = 0;
			if (refh_row(mgi, m1, m2, m3, c, un1, ui1, eps1) >= 0) {
				static const int slot[4] = {0, 6, 23, 40};
				gausmat2(1, mcmax1 * 4 + c, mcmax1 * 57 + slot[c], un1, ui1);
			} else {
				num0[mcmax1 * 4 + c] = 0;
			}
		}
	}
}

/* Thin wrappers around reference routines. */
void refh_ronurestrict(int mgi) { ronurestrict(mgi); }
void refh_vprestrict(int mgi) { vprestrict(mgi); }
void refh_vpprolong(int mgi) { vpprolong(mgi, 0); }
void refh_gausmat4(int am, int mgi)
{
	gausmat4(am, (mgp[mgi] + mgn[mgi]) * 4 - 1, mgp[mgi] * 4, mgi);
}

/* Value-mode stress/strain at one point (move3mg.c:3-710 with ynval=0):
 * kind 0..5 = xx,yy,zz,xy,xz,yz.  out[0]=sigma, out[1]=eps, out[2]=eta. */
void refh_stress(int kind, long m1, long m2, long m3, double *out)
{
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	double s = 0;
	un1[0] = 0;
	switch (kind) {
	case 0: s = sxxcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	case 1: s = syycalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	case 2: s = szzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	case 3: s = sxycalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	case 4: s = sxzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	default: s = syzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); break;
	}
	out[0] = s; out[1] = eps1[0]; out[2] = eps1[2];
}

/* Full reference Stokes solve (vpiterate, move3mg.c:1585-2636) from the
 * current fields; used for the CPU baseline timing.  Returns wall seconds. */
double refh_vpiterate(int nthreads)
{
	double t0;
	if (nthreads > 0) omp_set_num_threads(nthreads);
	fl0cyc[0][2] = 1000000;
	printmod = 0;
	t0 = omp_get_wtime();
	vpiterate(0, 0);
	return omp_get_wtime() - t0;
}

/* Marker -> node transfer and RK4 advection through the reference. */
void refh_ronurecalc(void) { ronurecalc(); }
void refh_movemark(void) { movemark(); }
long refh_m1serch(double x) { return m1serch(x); }
long refh_m2serch(double y) { return m2serch(y); }
long refh_m3serch(double z) { return m3serch(z); }
void refh_vxyzcalc1(double x, double y, double z, double *out)
{
	long wn1[100];
	vxyzcalc1(x, y, z, out, wn1);
}

/* The reference's scatter loops race between OpenMP threads
 * (move3mg.c:2662, :3185; mark3mg.c:460); golden vectors use one thread. */
void refh_set_threads(int n) { omp_set_num_threads(n); }

/* Value-mode strain rates and stresses of level 0 into the reference's own
 * sxx..syz / exx..eyz arrays, the loop at the end of vpiterate
 * (move3mg.c:1691-1764), from the current vx/vy/vz. */
void refh_fill_stress(void)
{
	long m1, m2, m3, n;
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	for (m3 = 0; m3 < mgz[0]; m3++)
	for (m1 = 0; m1 < mgx[0]; m1++)
	for (m2 = 0; m2 < mgy[0]; m2++) {
		n = m3 * mgxy[0] + m1 * mgy[0] + m2;
		if (m1 > 0 && m2 > 0 && m3 > 0) {
			sxx[n] = sxxcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); exx[n] = eps1[0];
			syy[n] = syycalc(m1, m2, m3, 0, 0, eps1, un1, ui1); eyy[n] = eps1[0];
			szz[n] = szzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); ezz[n] = eps1[0];
		}
		if (m1 > 0 && m1 < mgx[0] - 1 && m2 > 0 && m2 < mgy[0] - 1 && m3 < mgz[0] - 1) {
			sxy[n] = sxycalc(m1, m2, m3, 0, 0, eps1, un1, ui1); exy[n] = eps1[0];
		}
		if (m1 > 0 && m1 < mgx[0] - 1 && m2 < mgy[0] - 1 && m3 > 0 && m3 < mgz[0] - 1) {
			sxz[n] = sxzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); exz[n] = eps1[0];
		}
		if (m1 < mgx[0] - 1 && m2 > 0 && m2 < mgy[0] - 1 && m3 > 0 && m3 < mgz[0] - 1) {
			syz[n] = syzcalc(m1, m2, m3, 0, 0, eps1, un1, ui1); eyz[n] = eps1[0];
		}
	}
}

/* Residual-mode rows of the norm loop of vpiterate (move3mg.c:2405-2560):
 * kind 0 continerr, 1..3 x/y/zstokserr, 4 bonderr of sol0 unknown
 * 4 n + comp (comp = m1 here).  out[0] = return value, out[1] = eps1[0]. */
void refh_err_row(int kind, long m1, long m2, long m3, double *out)
{
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	eps1[0] = 0;
	switch (kind) {
	case 0: out[0] = continerr(m1, m2, m3, 1, 0, eps1, un1, ui1); break;
	case 1: out[0] = xstokserr(m1, m2, m3, 1, 0, eps1, un1, ui1); break;
	case 2: out[0] = ystokserr(m1, m2, m3, 1, 0, eps1, un1, ui1); break;
	case 3: out[0] = zstokserr(m1, m2, m3, 1, 0, eps1, un1, ui1); break;
	default: out[0] = bonderr(m1, 1, eps1, un1, ui1); break;
	}
	out[1] = eps1[0];
}

/* heaterr in right-hand-side mode (ynerr=0, heat3mg.c:54-160) at one level-0
 * node: out[0] = the returned rate ival/(ro cp), out[1] = the source it
 * stores in tk1 (radiogenic + adiabatic + shear heating). */
void refh_heat_rhs(long m1, long m2, long m3, double *out)
{
	long un1[MAXPOS];
	double ui1[MAXPOS], eps1[100];
	out[0] = heaterr(m1, m2, m3, 0, 0, eps1, un1, ui1);
	out[1] = tk1[m3 * mgxy[0] + m1 * mgy[0] + m2];
}

/* epscalc1 (mark3mg.c:2535-2838) at a point: out[0..5] eps1[4..9],
 * out[6] eps1[10] (P), out[7..12] eps1[54..59], out[13] eps1[60]. */
void refh_epscalc1(double x, double y, double z, double *out)
{
	double eps1[100];
	long wn1[100];
	int q;
	epscalc1(x, y, z, 0, eps1, wn1);
	for (q = 0; q < 7; q++) out[q] = eps1[4 + q];
	for (q = 0; q < 7; q++) out[7 + q] = eps1[54 + q];
}

/* The reference's run-file routines (load3mg.c): loadconf reads file.t3c and
 * mode.t3c in the current directory (and writes stop.yn there); saver and
 * loader write / read the .prn state file named by path (saver also writes
 * file.t3c and reads stop.yn in the current directory). */
int refh_loadconf(void) { return loadconf(); }
void refh_saver(const char *path, int f0, int n0)
{
	strncpy(fl1out, path, sizeof(fl1out) - 1);
	saver(f0, n0);
}
void refh_loader(const char *path)
{
	strncpy(fl1in, path, sizeof(fl1in) - 1);
	loader();
}
