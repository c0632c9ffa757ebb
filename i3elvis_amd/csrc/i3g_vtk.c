/*
 * ParaView output of the node fields and of the marker composition in the
 * layout of the reference's writer (i3vis_output_paraview.c): VTK XML
 * RectilinearGrid files with raw appended data, little endian, 32-bit
 * length prefixes, and the PVD time collection.  Host C.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/i3g.h"

/* an appended block: 32-bit byte count, then the bytes */
static int put_block(FILE *f, const void *p, size_t bytes)
{
	int32_t len = (int32_t)bytes;
	return fwrite(&len, sizeof(len), 1, f) == 1 && fwrite(p, 1, bytes, f) == bytes;
}

/* header of a RectilinearGrid piece with its three coordinate arrays;
 * returns the byte offset after them */
static long grid_header(FILE *f, long ex, long ey, long ez, long M, long N, long P)
{
	long off = 0;
	fprintf(f, "<VTKFile type=\"RectilinearGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
	fprintf(f, "  <RectilinearGrid WholeExtent=\"0 %ld 0 %ld 0 %ld\">\n", ex, ey, ez);
	fprintf(f, "    <Piece Extent=\"0 %ld 0 %ld 0 %ld\">\n", ex, ey, ez);
	fprintf(f, "      <Coordinates>\n");
	fprintf(f, "        <DataArray Name=\"xcoor\" type=\"Float64\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%d\">\n",
		(int)off);
	fprintf(f, "        </DataArray>\n");
	off += 4 + M * 8;
	fprintf(f, "        <DataArray Name=\"ycoor\" type=\"Float64\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%d\">\n",
		(int)off);
	fprintf(f, "        </DataArray>\n");
	off += 4 + N * 8;
	fprintf(f, "        <DataArray Name=\"zcoor\" type=\"Float64\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%d\">\n",
		(int)off);
	fprintf(f, "        </DataArray>\n");
	fprintf(f, "      </Coordinates>\n");
	off += 4 + P * 8;
	return off;
}

int i3g_write_vtr_nodal(const char *path, long nx, long ny, long nz, const double *gx, const double *gy,
			const double *gz, int nfield, const char *const *names, const void *const *fields, const int *f64)
{
	if (!path || nx < 1 || ny < 1 || nz < 1 || !gx || !gy || !gz || nfield < 0 || (nfield && (!names || !fields)))
		return I3G_EINVAL;
	long L = nx * ny * nz;
	double bytes = (double)(nx + ny + nz) * 8 + 4.0 * (3 + nfield);
	for (int q = 0; q < nfield; q++) {
		if (!names[q] || !fields[q]) return I3G_EINVAL;
		bytes += (double)L * (f64 && f64[q] ? 8 : 4);
	}
	/* the appended offsets are 32-bit in the file */
	if (bytes > 2147483647.0) return I3G_ECAP;
	FILE *f = fopen(path, "wb");
	if (!f) return I3G_EINVAL;
	long off = grid_header(f, nx - 1, ny - 1, nz - 1, nx, ny, nz);
	fprintf(f, "      <PointData>\n");
	for (int q = 0; q < nfield; q++) {
		int d = f64 && f64[q];
		fprintf(f, "        <DataArray Name=\"%s\" type=\"%s\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%d\">\n",
			names[q], d ? "Float64" : "Float32", (int)off);
		fprintf(f, "        </DataArray>\n");
		off += 4 + L * (d ? 8 : 4);
	}
	fprintf(f, "      </PointData>\n");
	fprintf(f, "    </Piece>\n");
	fprintf(f, "  </RectilinearGrid>\n");
	fprintf(f, "  <AppendedData encoding=\"raw\">\n");
	fprintf(f, "_");
	int ok = put_block(f, gx, nx * 8) && put_block(f, gy, ny * 8) && put_block(f, gz, nz * 8);
	double *buf = ok && nfield ? (double *)malloc(sizeof(double) * L) : NULL;
	if (nfield && !buf) ok = 0;
	for (int q = 0; q < nfield && ok; q++) {
		/* node m3*nx*ny + m1*ny + m2 (y fastest) -> m1 + m2*nx + m3*nx*ny
		 * (x fastest), as ivis_2_pv_permute_nodal_array_F32/F64 */
		if (f64 && f64[q]) {
			const double *src = (const double *)fields[q];
			for (long k = 0; k < nz; k++)
				for (long j = 0; j < ny; j++)
					for (long i = 0; i < nx; i++) buf[i + j * nx + k * nx * ny] = src[k * nx * ny + i * ny + j];
			ok = put_block(f, buf, L * 8);
		} else {
			const float *src = (const float *)fields[q];
			float *fb = (float *)buf;
			for (long k = 0; k < nz; k++)
				for (long j = 0; j < ny; j++)
					for (long i = 0; i < nx; i++) fb[i + j * nx + k * nx * ny] = src[k * nx * ny + i * ny + j];
			ok = put_block(f, fb, L * 4);
		}
	}
	free(buf);
	fprintf(f, "  </AppendedData>\n");
	fprintf(f, "</VTKFile>\n");
	if (fclose(f) != 0) ok = 0;
	return ok ? I3G_OK : I3G_EINVAL;
}

/* i3vis_generate_compo_map (i3vis_output_paraview.c:1111-1314): every cell
 * of the uniform mx x my x mz grid takes a marker type.  A marker of type
 * 0..50 inside the domain claims its cell when the cell is already claimed,
 * or when it lies closer (in x and y) to the cell centre than the cell's
 * recorded marker, which starts as marker 0; unclaimed cells then take, in
 * cell order, the type of the closest (x, y) recorded marker of their six
 * neighbours that are claimed by then (including cells filled earlier in
 * this pass, whose recorded marker stays marker 0). */
static void compo_map(double Lx, double Ly, double Lz, long mx, long my, long mz, long np, const float *xp,
		      const float *yp, const float *zp, const signed char *rt, int32_t *compo, long *cm)
{
	long ncells = mx * my * mz;
	double dx = Lx / (double)mx, dy = Ly / (double)my, dz = Lz / (double)mz;
	for (long c = 0; c < ncells; c++) {
		compo[c] = -1;
		cm[c] = 0;
	}
	for (long p = 0; p < np; p++) {
		if (rt[p] < 0 || rt[p] > 50) continue;
		if (xp[p] < 0.0 || yp[p] < 0.0 || zp[p] < 0.0 || xp[p] > Lx || yp[p] > Ly || zp[p] > Lz) continue;
		long ii = (long)(xp[p] / dx), jj = (long)(yp[p] / dy), kk = (long)(zp[p] / dz);
		if (ii == mx) ii--;
		if (jj == my) jj--;
		if (kk == mz) kk--;
		long cidx = ii + jj * mx + kk * mx * my;
		if (compo[cidx] != -1) {
			cm[cidx] = p;
			compo[cidx] = rt[p];
		} else {
			long q = cm[cidx];
			double cx = ii * dx + 0.5 * dx, cy = jj * dy + 0.5 * dy;
			double s1 = (xp[q] - cx) * (xp[q] - cx) + (yp[q] - cy) * (yp[q] - cy);
			double s2 = (xp[p] - cx) * (xp[p] - cx) + (yp[p] - cy) * (yp[p] - cy);
			if (s2 < s1) {
				cm[cidx] = p;
				compo[cidx] = rt[p];
			}
		}
	}
	for (long kk = 0; kk < mz; kk++)
		for (long jj = 0; jj < my; jj++)
			for (long ii = 0; ii < mx; ii++) {
				long cidx = ii + jj * mx + kk * mx * my;
				if (compo[cidx] != -1) continue;
				double cx = ii * dx + 0.5 * dx, cy = jj * dy + 0.5 * dy;
				double maxsep = 1.0e32;
				long nb[6];
				int ok[6] = {ii - 1 >= 0, ii + 1 < mx, jj - 1 >= 0, jj + 1 < my, kk - 1 >= 0, kk + 1 < mz};
				nb[0] = cidx - 1; nb[1] = cidx + 1; nb[2] = cidx - mx; nb[3] = cidx + mx;
				nb[4] = cidx - mx * my; nb[5] = cidx + mx * my;
				for (int q = 0; q < 6; q++) {
					if (!ok[q] || compo[nb[q]] == -1) continue;
					long cp = cm[nb[q]];
					double sep = (xp[cp] - cx) * (xp[cp] - cx) + (yp[cp] - cy) * (yp[cp] - cy);
					if (sep < maxsep) {
						maxsep = sep;
						compo[cidx] = rt[cp];
					}
				}
			}
}

int i3g_write_vtr_composition(const char *path, double xsize, double ysize, double zsize, long mx, long my, long mz,
			      long n, const float *x, const float *y, const float *z, const signed char *type)
{
	if (!path || mx < 1 || my < 1 || mz < 1 || n < 1 || !x || !y || !z || !type) return I3G_EINVAL;
	long M = mx + 1, N = my + 1, P = mz + 1, C = mx * my * mz;
	if ((double)(M + N + P) * 8 + (double)C * 4 + 16.0 > 2147483647.0) return I3G_ECAP;
	double *g = (double *)malloc(sizeof(double) * (M + N + P));
	int32_t *compo = (int32_t *)malloc(sizeof(int32_t) * C);
	long *cm = (long *)malloc(sizeof(long) * C);
	if (!g || !compo || !cm) {
		free(g); free(compo); free(cm);
		return I3G_ENOMEM;
	}
	double dx = xsize / (double)mx, dy = ysize / (double)my, dz = zsize / (double)mz;
	for (long i = 0; i < M; i++) g[i] = i * dx;
	for (long j = 0; j < N; j++) g[M + j] = j * dy;
	for (long k = 0; k < P; k++) g[M + N + k] = k * dz;
	compo_map(xsize, ysize, zsize, mx, my, mz, n, x, y, z, type, compo, cm);
	free(cm);
	FILE *f = fopen(path, "wb");
	if (!f) {
		free(g); free(compo);
		return I3G_EINVAL;
	}
	long off = grid_header(f, mx, my, mz, M, N, P);
	fprintf(f, "      <PointData>\n");
	fprintf(f, "      </PointData>\n");
	fprintf(f, "      <CellData>\n");
	fprintf(f, "        <DataArray Name=\"%s\" type=\"Int32\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%d\">\n",
		"composition", (int)off);
	fprintf(f, "        </DataArray>\n");
	fprintf(f, "      </CellData>\n");
	fprintf(f, "    </Piece>\n");
	fprintf(f, "  </RectilinearGrid>\n");
	fprintf(f, "  <AppendedData encoding=\"raw\">\n");
	fprintf(f, "_");
	int ok = put_block(f, g, M * 8) && put_block(f, g + M, N * 8) && put_block(f, g + M + N, P * 8) &&
		 put_block(f, compo, C * 4);
	fprintf(f, "  </AppendedData>\n");
	fprintf(f, "</VTKFile>\n");
	if (fclose(f) != 0) ok = 0;
	free(g);
	free(compo);
	return ok ? I3G_OK : I3G_EINVAL;
}

int i3g_pvd_open(const char *path)
{
	if (!path) return I3G_EINVAL;
	FILE *f = fopen(path, "w");
	if (!f) return I3G_EINVAL;
	fprintf(f, "<?xml version=\"1.0\"?>\n");
	fprintf(f, "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
	fprintf(f, "<Collection>\n");
	fprintf(f, "</Collection>\n");
	fprintf(f, "</VTKFile>\n");
	return fclose(f) == 0 ? I3G_OK : I3G_EINVAL;
}

/* the lines before "</Collection>" kept, one DataSet line added, the
 * closing tags rewritten (ParaviewPVDAppend) */
int i3g_pvd_append(const char *path, double time, const char *datafile, const char *dir)
{
	if (!path || !datafile) return I3G_EINVAL;
	FILE *f = fopen(path, "r");
	if (!f) return I3G_EINVAL;
	size_t cap = 1 << 16, len = 0;
	char *keep = (char *)malloc(cap), line[10000];
	if (!keep) { fclose(f); return I3G_ENOMEM; }
	keep[0] = 0;
	while (fgets(line, sizeof(line) - 1, f)) {
		if (strncmp(line, "</Collection>", 13) == 0) break;
		size_t l = strlen(line);
		if (len + l + 1 > cap) {
			while (len + l + 1 > cap) cap *= 2;
			char *k2 = (char *)realloc(keep, cap);
			if (!k2) { free(keep); fclose(f); return I3G_ENOMEM; }
			keep = k2;
		}
		memcpy(keep + len, line, l + 1);
		len += l;
	}
	fclose(f);
	f = fopen(path, "w");
	if (!f) { free(keep); return I3G_EINVAL; }
	fputs(keep, f);
	if (dir) fprintf(f, "  <DataSet timestep=\"%1.6e\" file=\"%s/%s\"/>\n", time, dir, datafile);
	else fprintf(f, "  <DataSet timestep=\"%1.6e\" file=\"./%s\"/>\n", time, datafile);
	fprintf(f, "</Collection>\n");
	fprintf(f, "</VTKFile>\n");
	free(keep);
	return fclose(f) == 0 ? I3G_OK : I3G_EINVAL;
}
