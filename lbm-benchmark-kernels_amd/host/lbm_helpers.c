/*
 * lbm_helpers.c -- host-side helpers of the reference (src/Kernel.c),
 * restated over the plugin ABI: every access goes through kd->GetNode /
 * kd->SetNode, so the helpers work for any kernel's node view.
 */
#include "lbm_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Direction table in the reference's order (src/Kernel.h:121-140). */
int D3Q19_X[D3Q19_NDIR]   = { 0, 0, 1,-1, 1, 1,-1,-1, 0, 0, 1,-1, 0, 0, 0, 0,-1, 1, 0 };
int D3Q19_Y[D3Q19_NDIR]   = { 1,-1, 0, 0, 1,-1, 1,-1, 0, 1, 0, 0,-1, 0,-1, 1, 0, 0, 0 };
int D3Q19_Z[D3Q19_NDIR]   = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1,-1,-1,-1,-1,-1, 0 };
int D3Q19_INV[D3Q19_NDIR] = { 1, 0, 3, 2, 7, 6, 5, 4,13,14,16,17,15, 8, 9,12,10,11,18 };
const char * D3Q19_NAMES[D3Q19_NDIR] = {
	"N", "S", "E", "W", "NE", "SE", "NW", "SW", "T", "TN", "TE", "TW", "TS",
	"B", "BS", "BN", "BW", "BE", "C" };

static inline int cell_index(const int d[3], int x, int y, int z)
{
	return (z * d[1] + y) * d[0] + x;     /* x fastest, src/Lattice.h:75 */
}

static inline int is_solid(const LatticeDesc * ld, int x, int y, int z)
{
	return ld->Lattice[cell_index(ld->Dims, x, y, z)] == LBM_CELL_OBSTACLE;
}

/* x momentum of a node view, summed left to right as src/Kernel.c:488-489. */
static inline PdfT ux_of(const PdfT * p)
{
	return p[D3Q19_DIR_E] + p[D3Q19_DIR_NE] + p[D3Q19_DIR_SE] + p[D3Q19_DIR_TE] + p[D3Q19_DIR_BE]
	     - p[D3Q19_DIR_W] - p[D3Q19_DIR_NW] - p[D3Q19_DIR_SW] - p[D3Q19_DIR_TW] - p[D3Q19_DIR_BW];
}

PdfT KernelDensity(KernelData * kd, LatticeDesc * ld)
{
	PdfT p[D3Q19_NDIR];
	PdfT total = 0.0;
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < ld->Dims[0]; ++x) {
		if (is_solid(ld, x, y, z)) continue;
		kd->GetNode(kd, x, y, z, p);
		PdfT node = 0.0;
		for (int d = 0; d < D3Q19_NDIR; ++d) node += p[d];
		total += node;
	}
	return total / ld->nFluid;
}

/* Rest-state equilibrium of the reference's initialisers
 * (src/Kernel.c:296-333): f_d = 3 w_d omega * (rho/3 - |u|^2/2 +- c.u +
 * 3/2 (c.u)^2) with u = 0; the velocity terms are kept so the rounding
 * matches term for term. */
static void rest_equilibrium(PdfT rho, PdfT omega, PdfT * f)
{
	const PdfT ux = 0.0, uy = 0.0, uz = 0.0;
	const PdfT ow0 = 3.0 * (1.0 / 3.0) * omega;
	const PdfT ow1 = 3.0 * (1.0 / 18.0) * omega;
	const PdfT ow2 = 3.0 * (1.0 / 36.0) * omega;
	const PdfT t = (1.0 / 3.0) * rho - 0.5 * (ux * ux + uy * uy + uz * uz);
	const PdfT a = ux - uy, b = ux + uy, c = ux - uz, e = ux + uz, g = uy - uz, h = uy + uz;

	f[D3Q19_DIR_C]  = ow0 * (t);
	f[D3Q19_DIR_NW] = ow2 * (t - a + 1.5 * (a * a));
	f[D3Q19_DIR_SE] = ow2 * (t + a + 1.5 * (a * a));
	f[D3Q19_DIR_NE] = ow2 * (t + b + 1.5 * (b * b));
	f[D3Q19_DIR_SW] = ow2 * (t - b + 1.5 * (b * b));
	f[D3Q19_DIR_TW] = ow2 * (t - c + 1.5 * (c * c));
	f[D3Q19_DIR_BE] = ow2 * (t + c + 1.5 * (c * c));
	f[D3Q19_DIR_TE] = ow2 * (t + e + 1.5 * (e * e));
	f[D3Q19_DIR_BW] = ow2 * (t - e + 1.5 * (e * e));
	f[D3Q19_DIR_TS] = ow2 * (t - g + 1.5 * (g * g));
	f[D3Q19_DIR_BN] = ow2 * (t + g + 1.5 * (g * g));
	f[D3Q19_DIR_TN] = ow2 * (t + h + 1.5 * (h * h));
	f[D3Q19_DIR_BS] = ow2 * (t - h + 1.5 * (h * h));
	f[D3Q19_DIR_N]  = ow1 * (t + uy + 1.5 * (uy * uy));
	f[D3Q19_DIR_S]  = ow1 * (t - uy + 1.5 * (uy * uy));
	f[D3Q19_DIR_E]  = ow1 * (t + ux + 1.5 * (ux * ux));
	f[D3Q19_DIR_W]  = ow1 * (t - ux + 1.5 * (ux * ux));
	f[D3Q19_DIR_T]  = ow1 * (t + uz + 1.5 * (uz * uz));
	f[D3Q19_DIR_B]  = ow1 * (t - uz + 1.5 * (uz * uz));
}

void KernelSetInitialDensity(LatticeDesc * ld, KernelData * kd, CaseData * cd)
{
	const int nX = ld->Dims[0];
	#pragma omp parallel for collapse(3) schedule(static)
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < nX; ++x) {
		if (is_solid(ld, x, y, z)) continue;
		PdfT f[D3Q19_NDIR];
		/* linear density ramp from inlet to outlet (src/Kernel.c:296) */
		const PdfT rho = cd->RhoIn + (cd->RhoOut - cd->RhoIn) * (x) / (nX - 1.0);
		rest_equilibrium(rho, cd->Omega, f);
		kd->SetNode(kd, x, y, z, f);
	}
}

void KernelSetInitialVelocity(LatticeDesc * ld, KernelData * kd, CaseData * cd)
{
	#pragma omp parallel for collapse(3) schedule(static)
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < ld->Dims[0]; ++x) {
		if (ld->Lattice[cell_index(ld->Dims, x, y, z)] != LBM_CELL_FLUID) continue;
		PdfT f[D3Q19_NDIR];
		/* the reference overwrites fluid cells with density 1 at rest
		 * (src/Kernel.c:350-434) */
		rest_equilibrium(1.0, cd->Omega, f);
		kd->SetNode(kd, x, y, z, f);
	}
}

void KernelAddBodyForce(KernelData * kd, LatticeDesc * ld, CaseData * cd)
{
	static const PdfT w[D3Q19_NDIR] = {
		1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
		1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
		1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0 };
	const PdfT F = cd->XForce;
	#pragma omp parallel for collapse(3) schedule(static)
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < ld->Dims[0]; ++x) {
		if (is_solid(ld, x, y, z)) continue;
		PdfT p[D3Q19_NDIR];
		kd->GetNode(kd, x, y, z, p);
		for (int d = 0; d < D3Q19_NDIR; ++d) p[d] = p[d] + 3.0 * w[d] * D3Q19_X[d] * F;   /* Luo, src/Kernel.c:795 */
		kd->SetNode(kd, x, y, z, p);
	}
}

void KernelVerifiy(LatticeDesc * ld, KernelData * kd, CaseData * cd, PdfT * errorNorm)
{
	const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
	const PdfT visc = (1.0 / cd->Omega - 0.5) / 3.0;
	PdfT * slice = (PdfT *)malloc(sizeof(PdfT) * (size_t)nY * nZ);
	if (!slice) { printf("ERROR: allocating %zu bytes failed.\n", sizeof(PdfT) * (size_t)nY * nZ); exit(1); }

	/* x velocity at x = nX/2, half-force corrected (src/Kernel.c:456-503) */
	const int xs = nX / 2;
	PdfT p[D3Q19_NDIR];
	for (int z = 0; z < nZ; ++z)
	for (int y = 0; y < nY; ++y) {
		if (is_solid(ld, xs, y, z)) { slice[y * nZ + z] = 0.0; continue; }
		kd->GetNode(kd, xs, y, z, p);
		PdfT u = ux_of(p);
		u += 0.5 * cd->XForce;
		slice[y * nZ + z] = u;
	}

	printf("# Kernel validation: writing profile to flow-profile.dat\n");
	FILE * fh = fopen("flow-profile.dat", "w");
	if (!fh) { printf("ERROR: opening file flow-profile.dat failed.\n"); exit(1); }
	fprintf(fh, "# Flow profile in Z direction. Taken at the middle of the X length (= %d) of total length %d.\n", nZ / 2, nZ);
	/* the two plotting hints of the reference's file (src/Kernel.c:561-562) */
	const char * name = "flow-profile.dat";
	fprintf(fh, "# Plot on terminal: gnuplot -e \"set terminal dumb; plot \\\"%s\\\" u 1:3 t \\\"analytical\\\", "
	            "\\\"\\\" u 1:4 t \\\"simulation\\\";\"\n", name);
	fprintf(fh, "# Plot graphically: gnuplot -e \"plot \\\"%s\\\" u 1:3 w linesp t \\\"analytical\\\", "
	            "\\\"\\\" u 1:4 w linesp t \\\"simulation\\\"; pause -1;\"\n", name);
	fprintf(fh, "# z coord., radius, analytic, simulation, diff abs, diff rel, undim_analytic, undim_sim\n");

	/* plane Poiseuille between walls half a cell inside the first and last
	 * layer; centre and radius taken from nY as src/Kernel.c:567-589 does */
	const int evenY = (nY % 2 == 0);
	const int yRow = (nY - evenY - 1) / 2;
	const PdfT center = nY / 2.0;
	const PdfT r2max = ((PdfT)nY / 2.0 - 1.0) * ((PdfT)nY / 2.0 - 1.0);
	const PdfT uMax = cd->XForce * r2max / (2.0 * visc);
	PdfT dev = 0.0;
	for (int z = 0; z < nZ; ++z) {
		const PdfT r = z - center + 0.5;
		const PdfT r2 = r * r;
		const PdfT ana = (r2 >= r2max) ? 0.0 : cd->XForce * (r2max - r2) / (2.0 * visc);
		const PdfT sim = evenY ? (slice[yRow * nZ + z] + slice[(yRow + 1) * nZ + z]) / 2.0 : slice[yRow * nZ + z];
		fprintf(fh, "%d\t%e\t%e\t%e\t%e\t", z, r / center, ana, sim, fabs(ana - sim));
		if (ana != 0.0) {
			const PdfT rel = (PdfT)fabs(ana - sim) / ana;
			fprintf(fh, "%e\t", rel);
			dev += rel * rel;
		} else {
			fprintf(fh, "0.0\t");
		}
		fprintf(fh, "%e\t%e\t\n", ana / uMax, sim / uMax);
	}
	fclose(fh);
	free(slice);
	*errorNorm = (PdfT)sqrt(dev);
	printf("# Kernel validation: L2 error norm of relative error: %e\n", *errorNorm);
}

void KernelStatistics(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration)
{
	KernelStatisticsAdv(kd, ld, cd, iteration, 0);
}

void KernelStatisticsAdv(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration, int forceOutput)
{
	if (iteration % cd->StatisticsModulus == 0 || forceOutput)
		printf("# iter: %4d   avg density: %e\n", iteration, KernelDensity(kd, ld));
	if (iteration % 10 != 0 && !forceOutput) return;

	const int nX = ld->Dims[0], nY = ld->Dims[1], nZ = ld->Dims[2];
	PdfT p[D3Q19_NDIR];

	/* mean x velocity over the mid-x cross section, appended per call */
	const int xm = nX / 2;
	double uSum = 0.0;
	int cnt = 0;
	for (int y = 0; y < nY; ++y)
	for (int z = 0; z < nZ; ++z) {
		if (is_solid(ld, xm, y, z)) continue;
		kd->GetNode(kd, xm, y, z, p);
		uSum += ux_of(p);
		++cnt;
	}
	FILE * fh = fopen("ux-progress.dat", iteration > 0 ? "a" : "w");
	if (!fh) { printf("ERROR: opening file ux-progress.dat failed.\n"); exit(1); }
	if (iteration == 0) {
		fprintf(fh, "# Average velocity in x direction of cross section in the middle (x = %d) of the geometry (NX = %d).\n", xm, nX);
		fprintf(fh, "# Plot on terminal: gnuplot -e \"set terminal dumb; plot \\\"%s\\\";\"\n", "ux-progress.dat");
		fprintf(fh, "# iteration, avg ux\n");
	}
	fprintf(fh, "%d %e\n", iteration, uSum / cnt);
	fclose(fh);

	/* density and x velocity averaged over every x cross section */
	fh = fopen("density-ux.dat", "w");
	if (!fh) { printf("ERROR: opening file density-ux.dat failed.\n"); exit(1); }
	fprintf(fh, "# Average density and average x velocity over each cross section in x direction. Snapshot taken at iteration %d.\n", iteration);
	fprintf(fh, "# Plot on terminal: gnuplot -e \"set terminal dumb; plot \\\"%s\\\" u 1:2; plot \\\"%s\\\" u 1:3;\"\n",
	        "density-ux.dat", "density-ux.dat");
	fprintf(fh, "# x, avg density, avg ux\n");
	for (int x = 0; x < nX; ++x) {
		double rs = 0.0, us = 0.0;
		int n = 0;
		for (int y = 0; y < nY; ++y)
		for (int z = 0; z < nZ; ++z) {
			if (is_solid(ld, x, y, z)) continue;
			kd->GetNode(kd, x, y, z, p);
			/* summation order of src/Kernel.c:732-737 */
			const PdfT rho = p[D3Q19_DIR_C]
				+ p[D3Q19_DIR_N] + p[D3Q19_DIR_E] + p[D3Q19_DIR_S] + p[D3Q19_DIR_W]
				+ p[D3Q19_DIR_NE] + p[D3Q19_DIR_SE] + p[D3Q19_DIR_SW] + p[D3Q19_DIR_NW]
				+ p[D3Q19_DIR_T] + p[D3Q19_DIR_TN] + p[D3Q19_DIR_TE] + p[D3Q19_DIR_TS] + p[D3Q19_DIR_TW]
				+ p[D3Q19_DIR_B] + p[D3Q19_DIR_BN] + p[D3Q19_DIR_BE] + p[D3Q19_DIR_BS] + p[D3Q19_DIR_BW];
			rs += rho;
			us += ux_of(p);
			++n;
		}
		fprintf(fh, "%d  %e  %e\n", x, rs / n, us / n);
	}
	fclose(fh);
}
