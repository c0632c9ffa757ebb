/*
 * i3g input/output of the reference's run files (host C, no device):
 *   the mode.t3c / file.t3c parameter deck (loadconf, load3mg.c:2-180) and
 *   the binary .prn state file (loader / saver, load3mg.c:704-1096).
 * The formats are the reference's, so a run can start from a reference
 * deck and state file and hand its state back to the reference.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/i3g.h"

/* ---- parameter deck -------------------------------------------------------- */

/* Next token of the deck: whitespace-separated, trailing characters <= 32
 * stripped, empty tokens and tokens starting with '/' (comments) skipped,
 * as ffscanf (load3mg.c:1102-1127).  0 at end of file. */
static int deck_token(FILE *f, char *sa, size_t cap)
{
	char fmt[16];
	snprintf(fmt, sizeof(fmt), "%%%zus", cap - 1);
	for (;;) {
		if (fscanf(f, fmt, sa) < 1) return 0;
		for (long q = (long)strlen(sa) - 1; q >= 0; q--) {
			if ((unsigned char)sa[q] <= 32) sa[q] = 0;
			else break;
		}
		if (strlen(sa) >= 1 && sa[0] != '/') return 1;
	}
}

#define TOK()                                                                  \
	do {                                                                   \
		if (!deck_token(f, sa, sizeof(sa))) { fclose(f); return I3G_EINVAL; } \
	} while (0)

int i3g_mode_read(const char *mode_path, const char *file_path, i3g_mode *m)
{
	char sa[256];
	FILE *f;
	int a, b;
	if (!mode_path || !m || m->fl0cap < 0 || (m->fl0cap > 0 && (!m->fl0out || !m->fl0otp || !m->fl0cyc || !m->fl0stp)))
		return I3G_EINVAL;
	m->fln3 = -1;
	if (file_path) {
		f = fopen(file_path, "rt");
		if (!f) return I3G_EINVAL;
		TOK();
		m->fln3 = atoi(sa) - 1;
		fclose(f);
	}
	f = fopen(mode_path, "rt");
	if (!f) return I3G_EINVAL;
	TOK();
	snprintf(m->fl1in, sizeof(m->fl1in), "%s", sa);
	TOK();
	m->fl1itp = sa[0] == 'b';
	m->fl0num = 0;
	TOK();
	while (sa[0] != '~') {
		if (m->fl0num >= m->fl0cap) { fclose(f); return I3G_ECAP; }
		int k = m->fl0num;
		snprintf(m->fl0out[k], sizeof(m->fl0out[k]), "%s", sa);
		TOK(); m->fl0otp[k] = sa[0] == 'b';
		TOK(); m->fl0cyc[k][0] = atoi(sa);
		TOK(); m->fl0stp[k][0] = atof(sa);
		TOK(); m->fl0stp[k][1] = atof(sa);
		TOK(); m->fl0stp[k][2] = atof(sa) * 3.15576e+7;
		for (a = 3; a <= 10; a++) { TOK(); m->fl0stp[k][a] = atof(sa); }
		for (a = 1; a <= 3; a++) { TOK(); m->fl0cyc[k][a] = atoi(sa); }
		m->fl0num++;
		TOK();
	}
	TOK(); m->printmod = atoi(sa);
	TOK(); m->movemod = atoi(sa);
	TOK(); m->tempmod = atoi(sa);
	TOK(); m->markmod = atoi(sa);
	TOK(); m->gridmod = atoi(sa);
	TOK(); m->outgrid = atoi(sa);
	TOK(); m->densimod = atoi(sa);
	TOK(); m->stp100 = atof(sa);
	TOK(); m->filesjob = atoi(sa);
	if (m->stp100 <= 0) { fclose(f); return I3G_EINVAL; }
	TOK(); m->eroslev = atof(sa);
	TOK(); m->sedilev = atof(sa);
	TOK(); m->waterlev = atof(sa);
	TOK(); m->cyc1max = atoi(sa);
	TOK(); m->DIVVMIN = atof(sa);
	TOK(); m->STOKSMIN = atof(sa);
	TOK(); m->DIVVMAX = atof(sa);
	TOK(); m->STOKSMAX = atof(sa);
	TOK(); m->multinum = atoi(sa);
	TOK(); m->multicyc = atoi(sa);
	if (m->multinum < 0 || m->multinum >= I3G_MAXLEV) { fclose(f); return I3G_ECAP; }
	for (a = 0; a <= 1; a++)
		for (b = 0; b <= m->multinum; b++) { TOK(); m->multinnn[a][b] = atoi(sa); }
	TOK(); m->p0koef = atof(sa);
	TOK(); m->p1koef = atof(sa);
	TOK(); m->p2koef = atof(sa);
	TOK(); m->v0koef = atof(sa);
	TOK(); m->v1koef = atof(sa);
	TOK(); m->nubeg = atof(sa);
	TOK(); m->nuend = atof(sa);
	TOK(); m->nukoef = atof(sa);
	TOK(); m->viscmod = atoi(sa);
	TOK(); m->spheryn = atoi(sa);
	TOK(); m->cyc2max = atoi(sa);
	TOK(); m->HEATMIN = atof(sa);
	TOK(); m->multinumt = atoi(sa);
	TOK(); m->multicyct = atoi(sa);
	if (m->multinumt < 0 || m->multinumt >= I3G_MAXLEV) { fclose(f); return I3G_ECAP; }
	for (a = 0; a <= 1; a++)
		for (b = 0; b <= m->multinumt; b++) { TOK(); m->multittt[a][b] = atoi(sa); }
	TOK(); m->t0koef = atof(sa);
	TOK(); m->t1koef = atof(sa);
	TOK(); m->heatdif = atof(sa);
	/* integers in the reference, read with atof */
	TOK(); m->frictyn = (int)atof(sa);
	TOK(); m->adiabyn = (int)atof(sa);
	{
		double *w[] = {&m->tkpor, &m->zmpor, &m->vyfluid, &m->vymelt, &m->dmwamin, &m->tdeep, &m->dtdeep,
			       &m->drdeep, &m->zdeep, &m->vdeep, &m->nudeep, &m->dxwater, &m->dywater, &m->dzwater,
			       &m->dserp, &m->yserp, &m->tserp, &m->lambmlt, &m->ydike, &m->dydike, &m->nusseltno,
			       &m->circdepth, &m->circtemp, &m->healrate, &m->meltmin, &m->maxwater, &m->maxpmelt,
			       &m->teclmin, &m->teclmax, &m->xvolc, &m->xtkplut};
		for (a = 0; a < (int)(sizeof(w) / sizeof(w[0])); a++) { TOK(); *w[a] = atof(sa); }
	}
	TOK(); m->timebeg = atof(sa) * 3.15576e+7;
	TOK(); m->timeend = atof(sa) * 3.15576e+7;
	TOK(); m->tempbeg = atof(sa);
	TOK(); m->convend = atof(sa) * 3.15576e+7;
	fclose(f);
	/* the data file of the run number in file.t3c */
	if (m->fln3 >= 0 && m->fln3 < m->fl0num) {
		snprintf(m->fl1in, sizeof(m->fl1in), "%s", m->fl0out[m->fln3]);
		m->fl1itp = m->fl0otp[m->fln3];
	} else {
		m->fln3 = -1;
	}
	return I3G_OK;
}

/* ---- .prn state file ------------------------------------------------------- */

#define PRN_NROCK 24  /* rock table columns in file order */
#define PRN_NNODE 16  /* node fields in file order */
#define PRN_NMARK 12  /* marker float fields in file order */

static int rd(FILE *f, void *p, size_t sz, size_t n)
{
	return n == 0 || fread(p, sz, n, f) == n;
}
static int wr(FILE *f, const void *p, size_t sz, size_t n)
{
	return n == 0 || fwrite(p, sz, n, f) == n;
}

static int prn_header(FILE *f, i3g_prn *p)
{
	unsigned char sz[4];
	if (!rd(f, sz, 1, 4)) return I3G_EINVAL;
	/* int, long, float, double sizes of the writer (load3mg.c:730-741) */
	if (sz[0] != sizeof(int) || sz[1] != sizeof(long) || sz[2] != sizeof(float) || sz[3] != sizeof(double))
		return I3G_EINVAL;
	long L[6];
	double D[3];
	if (!rd(f, L, sizeof(long), 6) || !rd(f, D, sizeof(double), 3)) return I3G_EINVAL;
	p->xnumx = L[0]; p->ynumy = L[1]; p->znumz = L[2];
	p->mnumx = L[3]; p->mnumy = L[4]; p->mnumz = L[5];
	p->xsize = D[0]; p->ysize = D[1]; p->zsize = D[2];
	if (!rd(f, &p->pxinit, sizeof(long), 1) || !rd(f, &p->pyinit, sizeof(long), 1) ||
	    !rd(f, &p->pzinit, sizeof(long), 1) || !rd(f, &p->pinit, sizeof(double), 1) ||
	    !rd(f, &p->gxkoef, sizeof(double), 1) || !rd(f, &p->gykoef, sizeof(double), 1) ||
	    !rd(f, &p->gzkoef, sizeof(double), 1) || !rd(f, &p->rocknum, sizeof(int), 1) ||
	    !rd(f, &p->bondnum, sizeof(long), 1) || !rd(f, &p->marknum, sizeof(long), 1) ||
	    !rd(f, &p->n0, sizeof(int), 1) || !rd(f, &p->timesum_years, sizeof(double), 1) ||
	    !rd(f, &p->gridcur, sizeof(double), 1) || !rd(f, &p->gridtot, sizeof(double), 1))
		return I3G_EINVAL;
	if (p->xnumx < 2 || p->ynumy < 2 || p->znumz < 2 || p->rocknum < 0 || p->bondnum < 0 || p->marknum < 0)
		return I3G_EINVAL;
	return I3G_OK;
}

int i3g_prn_read_header(const char *path, i3g_prn *p)
{
	if (!path || !p) return I3G_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return I3G_EINVAL;
	int rc = prn_header(f, p);
	fclose(f);
	return rc;
}

int i3g_prn_read(const char *path, i3g_prn *p)
{
	if (!path || !p || !p->rock || !p->node || !p->gx || !p->gy || !p->gz || !p->bond_at || !p->bond_rec ||
	    !p->bond_v || !p->bond_partner || !p->mark || !p->markt)
		return I3G_EINVAL;
	i3g_prn h = *p;
	FILE *f = fopen(path, "rb");
	if (!f) return I3G_EINVAL;
	int rc = prn_header(f, &h);
	if (rc) { fclose(f); return rc; }
	if (h.xnumx != p->xnumx || h.ynumy != p->ynumy || h.znumz != p->znumz || h.rocknum != p->rocknum ||
	    h.bondnum != p->bondnum || h.marknum != p->marknum) {
		fclose(f);
		return I3G_EINVAL; /* arrays sized from another header */
	}
	*p = h;
	long N = p->xnumx * p->ynumy * p->znumz, q, c;
	int ok = 1;
	/* rock table: 24 columns of rocknum doubles */
	for (c = 0; c < PRN_NROCK && ok; c++) ok = rd(f, p->rock + c * p->rocknum, sizeof(double), p->rocknum);
	/* node fields, interleaved per node in the file */
	for (q = 0; q < N && ok; q++) {
		float v[PRN_NNODE];
		ok = rd(f, v, sizeof(float), PRN_NNODE);
		for (c = 0; c < PRN_NNODE; c++) p->node[c * N + q] = v[c];
	}
	ok = ok && rd(f, p->gx, sizeof(float), p->xnumx) && rd(f, p->gy, sizeof(float), p->ynumy) &&
	     rd(f, p->gz, sizeof(float), p->znumz);
	for (q = 0; q < p->bondnum && ok; q++)
		ok = rd(f, &p->bond_at[q], sizeof(long), 1) && rd(f, &p->bond_rec[q], sizeof(long), 1) &&
		     rd(f, &p->bond_v[2 * q], sizeof(float), 2) && rd(f, &p->bond_partner[q], sizeof(long), 1);
	for (q = 0; q < p->marknum && ok; q++) {
		float v[PRN_NMARK];
		signed char t;
		ok = rd(f, v, sizeof(float), PRN_NMARK) && rd(f, &t, 1, 1);
		for (c = 0; c < PRN_NMARK; c++) p->mark[c * p->marknum + q] = v[c];
		p->markt[q] = t;
	}
	fclose(f);
	return ok ? I3G_OK : I3G_EINVAL;
}

int i3g_prn_write(const char *path, const i3g_prn *p)
{
	if (!path || !p || !p->rock || !p->node || !p->gx || !p->gy || !p->gz || (p->bondnum && (!p->bond_at ||
	    !p->bond_rec || !p->bond_v || !p->bond_partner)) || (p->marknum && (!p->mark || !p->markt)))
		return I3G_EINVAL;
	FILE *f = fopen(path, "wb");
	if (!f) return I3G_EINVAL;
	unsigned char sz[4] = {sizeof(int), sizeof(long), sizeof(float), sizeof(double)};
	long L[6] = {p->xnumx, p->ynumy, p->znumz, p->mnumx, p->mnumy, p->mnumz};
	double D[3] = {p->xsize, p->ysize, p->zsize};
	int ok = wr(f, sz, 1, 4) && wr(f, L, sizeof(long), 6) && wr(f, D, sizeof(double), 3) &&
		 wr(f, &p->pxinit, sizeof(long), 1) && wr(f, &p->pyinit, sizeof(long), 1) &&
		 wr(f, &p->pzinit, sizeof(long), 1) && wr(f, &p->pinit, sizeof(double), 1) &&
		 wr(f, &p->gxkoef, sizeof(double), 1) && wr(f, &p->gykoef, sizeof(double), 1) &&
		 wr(f, &p->gzkoef, sizeof(double), 1) && wr(f, &p->rocknum, sizeof(int), 1) &&
		 wr(f, &p->bondnum, sizeof(long), 1) && wr(f, &p->marknum, sizeof(long), 1) &&
		 wr(f, &p->n0, sizeof(int), 1) && wr(f, &p->timesum_years, sizeof(double), 1) &&
		 wr(f, &p->gridcur, sizeof(double), 1) && wr(f, &p->gridtot, sizeof(double), 1);
	long N = p->xnumx * p->ynumy * p->znumz, q, c;
	for (c = 0; c < PRN_NROCK && ok; c++) ok = wr(f, p->rock + c * p->rocknum, sizeof(double), p->rocknum);
	for (q = 0; q < N && ok; q++) {
		float v[PRN_NNODE];
		for (c = 0; c < PRN_NNODE; c++) v[c] = p->node[c * N + q];
		ok = wr(f, v, sizeof(float), PRN_NNODE);
	}
	ok = ok && wr(f, p->gx, sizeof(float), p->xnumx) && wr(f, p->gy, sizeof(float), p->ynumy) &&
	     wr(f, p->gz, sizeof(float), p->znumz);
	for (q = 0; q < p->bondnum && ok; q++)
		ok = wr(f, &p->bond_at[q], sizeof(long), 1) && wr(f, &p->bond_rec[q], sizeof(long), 1) &&
		     wr(f, &p->bond_v[2 * q], sizeof(float), 2) && wr(f, &p->bond_partner[q], sizeof(long), 1);
	for (q = 0; q < p->marknum && ok; q++) {
		float v[PRN_NMARK];
		for (c = 0; c < PRN_NMARK; c++) v[c] = p->mark[c * p->marknum + q];
		signed char t = p->markt[q];
		ok = wr(f, v, sizeof(float), PRN_NMARK) && wr(f, &t, 1, 1);
	}
	if (fclose(f) != 0) ok = 0;
	return ok ? I3G_OK : I3G_EINVAL;
}

/* ---- initial-condition boxes of a deck ------------------------------------- */

/* a coordinate of a rock box: 'm' metres scaled to a fraction of the size,
 * else a fraction (in3mg.c:573-575) */
static double box_frac(const char *sa, double size) { return sa[0] == 'm' ? atof(sa + 1) / size : atof(sa); }
/* a coordinate of a temperature box: 'm' metres, else a fraction of the
 * size (in3mg.c:811-825) */
static double box_metre(const char *sa, double size) { return sa[0] == 'm' ? atof(sa + 1) : atof(sa) * size; }

int i3g_deck_boxes(const char *path, double xsize, double ysize, double zsize, i3g_rock_box *rb, int maxrb, int *nrb,
		   i3g_temp_box *tb, int maxtb, int *ntb)
{
	if (!path || !rb || !tb || !nrb || !ntb || maxrb < 0 || maxtb < 0) return I3G_EINVAL;
	FILE *f = fopen(path, "r");
	if (!f) return I3G_EINVAL;
	char sa[256];
	*nrb = *ntb = 0;
	/* the rock table, the boundary conditions: two sections ended by '~' */
	int ends = 0;
	while (ends < 2) {
		TOK();
		if (sa[0] == '~') ends++;
	}
	/* rock bodies (in3mg.c:550-576) */
	TOK();
	while (sa[0] != '~') {
		i3g_rock_box b;
		memset(&b, 0, sizeof(b));
		if (sa[0] == 's' || sa[0] == 'c') {
			b.type = atoi(sa + 1);
			b.shape = sa[0] == 's' ? 1 : 2;
		} else {
			b.type = atoi(sa);
			b.shape = 0;
		}
		TOK(); b.ex = atof(sa);
		TOK(); b.e = atof(sa);
		for (int q = 0; q < 8; q++) {
			TOK(); b.x[q] = box_frac(sa, xsize);
			TOK(); b.y[q] = box_frac(sa, ysize);
			TOK(); b.z[q] = box_frac(sa, zsize);
		}
		if (*nrb >= maxrb) { fclose(f); return I3G_ECAP; }
		rb[(*nrb)++] = b;
		TOK();
	}
	/* temperature boxes (in3mg.c:806-829) */
	TOK();
	while (sa[0] != '~') {
		i3g_temp_box b;
		memset(&b, 0, sizeof(b));
		b.type = atoi(sa);
		TOK(); b.x0 = box_metre(sa, xsize);
		TOK(); b.x2 = box_metre(sa, xsize);
		TOK(); b.y0 = box_metre(sa, ysize);
		TOK(); b.y1 = box_metre(sa, ysize);
		TOK(); b.y2 = box_metre(sa, ysize);
		TOK(); b.y3 = box_metre(sa, ysize);
		TOK(); b.z0 = box_metre(sa, zsize);
		TOK(); b.x5 = box_metre(sa, xsize);
		TOK(); b.x7 = box_metre(sa, xsize);
		TOK(); b.y4 = box_metre(sa, ysize);
		TOK(); b.y5 = box_metre(sa, ysize);
		TOK(); b.y6 = box_metre(sa, ysize);
		TOK(); b.y7 = box_metre(sa, ysize);
		TOK(); b.z7 = box_metre(sa, zsize);
		for (int q = 0; q < 8; q++) {
			TOK(); b.t[q] = atof(sa);
		}
		if (*ntb >= maxtb) { fclose(f); return I3G_ECAP; }
		tb[(*ntb)++] = b;
		TOK();
	}
	fclose(f);
	return I3G_OK;
}
