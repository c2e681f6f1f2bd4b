/*
 * ref_harness_common.c -- glue shared by the list and aa harnesses.
 * TEST INFRASTRUCTURE ONLY (see ref_harness.h).  Compiled against the
 * reference headers by oracle/Makefile.
 */
#include "Kernel.h"
#include "Geometry.h"
#include "Padding.h"
#include "Memory.h"

#include "ref_harness.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* LatticeDesc over a caller-provided mask, counts as GeoCreateByType
 * computes them (src/Geometry.c:260-278). */
void refh_make_lattice(LatticeDesc * ld, const int * mask, const int dims[3], const int periodic[3])
{
	int n = dims[0] * dims[1] * dims[2];
	memset(ld, 0, sizeof(*ld));
	ld->Dims[0] = dims[0]; ld->Dims[1] = dims[1]; ld->Dims[2] = dims[2];
	ld->nCells = n;
	ld->PeriodicX = periodic[0]; ld->PeriodicY = periodic[1]; ld->PeriodicZ = periodic[2];
	ld->Name = "harness";
	MemAlloc((void **)&ld->Lattice, sizeof(LatticeT) * (size_t)n);
	for (int i = 0; i < n; ++i) {
		ld->Lattice[i] = mask[i];
		switch (mask[i]) {
			case LAT_CELL_OBSTACLE: ld->nObst++; break;
			case LAT_CELL_INLET:    ld->nInlet++;  ld->nFluid++; break;
			case LAT_CELL_OUTLET:   ld->nOutlet++; ld->nFluid++; break;
			default:                ld->nFluid++; break;
		}
	}
}

static int g_vtkOutput = 0, g_vtkModulus = 100, g_statisticsModulus = 100;

void ref_set_output(int vtkOutput, int vtkModulus, int statisticsModulus)
{
	g_vtkOutput = vtkOutput;
	g_vtkModulus = vtkModulus;
	g_statisticsModulus = statisticsModulus;
}

void refh_case(CaseData * cd, double omega, double xForce, int iterations)
{
	memset(cd, 0, sizeof(*cd));
	cd->Omega = omega;
	cd->RhoIn = 1.0;
	cd->RhoOut = 1.0;
	cd->XForce = xForce;
	cd->MaxIterations = iterations;
	cd->VtkOutput = g_vtkOutput;
	cd->VtkModulus = g_vtkModulus;
	cd->StatisticsModulus = g_statisticsModulus;
}

void refh_threads(int nThreads)
{
#ifdef _OPENMP
	omp_set_num_threads(nThreads > 0 ? nThreads : 1);
#else
	(void)nThreads;
#endif
}

static const PdfT refh_w[N_D3Q19] = {
	1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
	1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
	1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
	1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
	1.0 / 3.0 };

void refh_init(LatticeDesc * ld, KernelData * kd, CaseData * cd, int mode, double rho)
{
	if (mode == REF_INIT_VERIFY) {
		KernelSetInitialDensity(ld, kd, cd);
		KernelSetInitialVelocity(ld, kd, cd);
	}
	else if (mode == REF_INIT_WEIGHTS) {
		PdfT v[N_D3Q19];
		for (int d = 0; d < N_D3Q19; ++d) v[d] = refh_w[d] * rho;
		for (int z = 0; z < ld->Dims[2]; ++z)
		for (int y = 0; y < ld->Dims[1]; ++y)
		for (int x = 0; x < ld->Dims[0]; ++x)
			if (ld->Lattice[L_INDEX_4(ld->Dims, x, y, z)] != LAT_CELL_OBSTACLE)
				kd->SetNode(kd, x, y, z, v);
	}
}

void refh_finish(LatticeDesc * ld, KernelData * kd, CaseData * cd, int doVerify,
                 double * dump, ref_result_t * res)
{
	if (dump) {
		size_t k = 0;
		for (int z = 0; z < ld->Dims[2]; ++z)
		for (int y = 0; y < ld->Dims[1]; ++y)
		for (int x = 0; x < ld->Dims[0]; ++x)
			if (ld->Lattice[L_INDEX_4(ld->Dims, x, y, z)] != LAT_CELL_OBSTACLE) {
				kd->GetNode(kd, x, y, z, dump + k * N_D3Q19);
				++k;
			}
	}
	res->nFluid = ld->nFluid;
	res->duration = kd->Duration;
	res->density = KernelDensity(kd, ld);
	res->verifyNorm = -1.0;
	if (doVerify) {
		PdfT norm = -1.0;
		KernelVerifiy(ld, kd, cd, &norm);
		res->verifyNorm = norm;
	}
}

int ref_geometry(const char * type, const int dims[3], const int periodic[3],
                 int * mask, int ldPeriodic[3])
{
	int d[3] = { dims[0], dims[1], dims[2] };
	int p[3] = { periodic[0], periodic[1], periodic[2] };
	LatticeDesc ld;
	GeoCreateByStr(type, d, p, &ld);
	memcpy(mask, ld.Lattice, sizeof(int) * (size_t)d[0] * d[1] * d[2]);
	ldPeriodic[0] = ld.PeriodicX; ldPeriodic[1] = ld.PeriodicY; ldPeriodic[2] = ld.PeriodicZ;
	int nFluid = ld.nFluid;
	MemFree((void **)&ld.Lattice);
	return nFluid;
}

int ref_pad_cells(int nCells, const char * str)
{
	PadInfo * pi = str ? PadInfoFromStr(str) : NULL;
	int r = PadCells(nCells, sizeof(PdfT), &pi);
	PadInfoFree(pi);
	return r;
}
