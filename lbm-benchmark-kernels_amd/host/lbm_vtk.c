/*
 * lbm_vtk.c -- legacy VTK snapshot of a lattice (src/Vtk.c:51-244): the
 * reference's file name, layout and encoding, so that files of both
 * programs compare byte for byte.
 *
 * vtk/file-<iteration %04d>.vtk, "BINARY" structured points over the whole
 * box (x fastest): the cell types as unsigned char, the density as
 * big-endian double (0 in solid cells) and the velocity (sum c_d f_d, plus
 * half the body force in x as the reference's VERIFICATION build writes it)
 * as three big-endian doubles per cell (0 in solid cells).  Node values come
 * from kd->GetNode.
 */
#include "lbm_host.h"

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

static void put_be64(FILE * fh, double v)
{
	uint64_t u;
	memcpy(&u, &v, sizeof(u));
	u = __builtin_bswap64(u);
	fwrite(&u, sizeof(u), 1, fh);
}

/* the output directory, created on first use (src/Vtk.c:60-88) */
static void vtk_dir(void)
{
	struct stat st;
	if (stat("vtk", &st) == 0) {
		if (!S_ISDIR(st.st_mode)) {
			printf("ERROR: cannot create subdirectory vtk as already a file with the same name exists.\n");
			exit(1);
		}
		return;
	}
	if (mkdir("vtk", 0700) != 0) {
		printf("ERROR: cannot create directory vtk - %d: %s\n", errno, strerror(errno));
		exit(1);
	}
	printf("# created directory vtk.\n");
}

void VtkWrite(LatticeDesc * ld, KernelData * kd, CaseData * cd, int iteration)
{
	vtk_dir();
	char name[1024];
	snprintf(name, sizeof(name), "vtk/file-%04d.vtk", iteration);
	printf("# VTK: writing file %s\n", name);
	FILE * fh = fopen(name, "w");
	if (!fh) {
		printf("ERROR: opening file %s failed.\n", name);
		exit(1);
	}
	const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
	const size_t nCell = (size_t)nX * nY * nZ;
	fprintf(fh, "# vtk DataFile Version 1.0\n");
	fprintf(fh, "Comment: lid driven cavity, iteration % 4d\n", iteration);
	fprintf(fh, "BINARY\n");
	fprintf(fh, "DATASET STRUCTURED_POINTS\n");
	fprintf(fh, "DIMENSIONS %d %d %d\n", nX, nY, nZ);
	fprintf(fh, "ORIGIN 0 0 0 \n");
	fprintf(fh, "SPACING 1 1 1\n");
	fprintf(fh, "POINT_DATA %d\n", nX * nY * nZ);

	fprintf(fh, "SCALARS NodesTypes unsigned_char 1\n");
	fprintf(fh, "LOOKUP_TABLE default\n");
	for (size_t i = 0; i < nCell; ++i) {
		const unsigned char t = (unsigned char)ld->Lattice[i];
		fwrite(&t, 1, 1, fh);
	}

	PdfT p[D3Q19_NDIR];
	fprintf(fh, "SCALARS Density double\n");
	fprintf(fh, "LOOKUP_TABLE default\n");
	for (int z = 0; z < nZ; ++z)
	for (int y = 0; y < nY; ++y)
	for (int x = 0; x < nX; ++x) {
		double rho = 0.0;
		if (ld->Lattice[((size_t)z * nY + y) * nX + x] != LBM_CELL_OBSTACLE) {
			kd->GetNode(kd, x, y, z, p);
			for (int d = 0; d < D3Q19_NDIR; ++d) rho += p[d];     /* direction order */
		}
		put_be64(fh, rho);
	}

	fprintf(fh, "VECTORS VelocityVectors double\n");
	for (int z = 0; z < nZ; ++z)
	for (int y = 0; y < nY; ++y)
	for (int x = 0; x < nX; ++x) {
		double u[3] = {0.0, 0.0, 0.0};
		if (ld->Lattice[((size_t)z * nY + y) * nX + x] != LBM_CELL_OBSTACLE) {
			kd->GetNode(kd, x, y, z, p);
			/* left-to-right sums of src/Vtk.c:211-216 */
			u[0] = p[D3Q19_DIR_E] + p[D3Q19_DIR_NE] + p[D3Q19_DIR_SE] + p[D3Q19_DIR_TE] + p[D3Q19_DIR_BE]
			     - p[D3Q19_DIR_W] - p[D3Q19_DIR_NW] - p[D3Q19_DIR_SW] - p[D3Q19_DIR_TW] - p[D3Q19_DIR_BW];
			u[1] = p[D3Q19_DIR_N] + p[D3Q19_DIR_NE] + p[D3Q19_DIR_NW] + p[D3Q19_DIR_TN] + p[D3Q19_DIR_BN]
			     - p[D3Q19_DIR_S] - p[D3Q19_DIR_SE] - p[D3Q19_DIR_SW] - p[D3Q19_DIR_TS] - p[D3Q19_DIR_BS];
			u[2] = p[D3Q19_DIR_T] + p[D3Q19_DIR_TE] + p[D3Q19_DIR_TW] + p[D3Q19_DIR_TN] + p[D3Q19_DIR_TS]
			     - p[D3Q19_DIR_B] - p[D3Q19_DIR_BE] - p[D3Q19_DIR_BW] - p[D3Q19_DIR_BN] - p[D3Q19_DIR_BS];
			u[0] += 0.5 * cd->XForce;
		}
		put_be64(fh, u[0]);
		put_be64(fh, u[1]);
		put_be64(fh, u[2]);
	}
	fclose(fh);
}
