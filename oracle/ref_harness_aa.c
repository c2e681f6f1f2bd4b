/*
 * ref_harness_aa.c -- drives the reference aa-soa kernel.
 * TEST INFRASTRUCTURE ONLY (see ref_harness.h).  Compiled against the
 * reference headers with DATA_LAYOUT_SOA and PROP_MODEL_AA, the defines the
 * reference's own build uses for this kernel (src/Makefile:370-372).
 */
#include "BenchKernelD3Q19Aa.h"
#include "BenchKernelD3Q19AaCommon.h"
#include "Memory.h"

#include "ref_harness.h"

#include <string.h>

void refh_make_lattice(LatticeDesc * ld, const int * mask, const int dims[3], const int periodic[3]);
void refh_case(CaseData * cd, double omega, double xForce, int iterations);
void refh_threads(int nThreads);
void refh_init(LatticeDesc * ld, KernelData * kd, CaseData * cd, int mode, double rho);
void refh_finish(LatticeDesc * ld, KernelData * kd, CaseData * cd, int doVerify,
                 double * dump, ref_result_t * res);

int ref_aa_run(const int * mask, const int dims[3], const int periodic[3],
               int nThreads, int nKernelArgs, char ** kernelArgs,
               double omega, double xForce, int iterations, int initMode, double rho,
               int doVerify, double * dump, ref_result_t * res)
{
	LatticeDesc ld;
	CaseData cd;
	Parameters p;
	KernelData * kd = NULL;

	refh_threads(nThreads);
	refh_make_lattice(&ld, mask, dims, periodic);
	refh_case(&cd, omega, xForce, iterations);
	memset(&p, 0, sizeof(p));
	p.nKernelArgs = nKernelArgs;
	p.KernelArgs = kernelArgs;

	D3Q19AaInit_AaSoA(&ld, &kd, &p);

	/* The reference never initialises the ghost layer
	 * (src/BenchKernelD3Q19AaCommon.c:376-443); fresh pages read as 0.0.
	 * Zero it explicitly so runs do not depend on allocator history. */
	{
		int * g = kd->GlobalDims;
		size_t n = (size_t)N_D3Q19 * g[0] * g[1] * g[2];
		PdfT * pdf = kd->Pdfs[0];
		for (int d = 0; d < N_D3Q19; ++d)
		for (int x = 0; x < g[0]; ++x)
		for (int y = 0; y < g[1]; ++y)
		for (int z = 0; z < g[2]; ++z) {
			int ghost = x == 0 || y == 0 || z == 0 || x == g[0] - 1 || y == g[1] - 1 || z == g[2] - 1;
			if (ghost) pdf[P_INDEX_5(g, x, y, z, d)] = 0.0;
		}
		(void)n;
	}

	memset(res, 0, sizeof(*res));
	res->nCells = kd->GlobalDims[0] * kd->GlobalDims[1] * kd->GlobalDims[2];
	res->nBounceBack = kd->nBounceBackPdfs;

	refh_init(&ld, kd, &cd, initMode, rho);
	if (iterations > 0) kd->Kernel(&ld, kd, &cd);
	refh_finish(&ld, kd, &cd, doVerify, dump, res);

	D3Q19AaDeinit_AaSoA(&ld, &kd);
	MemFree((void **)&ld.Lattice);
	return 0;
}
