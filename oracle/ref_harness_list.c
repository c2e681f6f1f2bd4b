/*
 * ref_harness_list.c -- drives the reference list-aa-pv-soa kernel.
 * TEST INFRASTRUCTURE ONLY (see ref_harness.h).  Compiled against the
 * reference headers with DATA_LAYOUT_SOA and PROP_MODEL_PUSH, the defines the
 * reference's own build uses for this kernel (src/Makefile:342-344).
 */
#include "BenchKernelD3Q19ListAaPv.h"
#include "BenchKernelD3Q19ListAaPvCommon.h"
#include "Memory.h"

#include "ref_harness.h"

#include <string.h>

void refh_make_lattice(LatticeDesc * ld, const int * mask, const int dims[3], const int periodic[3]);
void refh_case(CaseData * cd, double omega, double xForce, int iterations);
void refh_threads(int nThreads);
void refh_init(LatticeDesc * ld, KernelData * kd, CaseData * cd, int mode, double rho);
void refh_finish(LatticeDesc * ld, KernelData * kd, CaseData * cd, int doVerify,
                 double * dump, ref_result_t * res);

int ref_list_run(const int * mask, const int dims[3], const int periodic[3],
                 int nThreads, int nKernelArgs, char ** kernelArgs,
                 double omega, double xForce, int iterations, int initMode, double rho,
                 int doVerify,
                 double * dump, uint32_t * adj, uint32_t * grid, uint32_t * coords,
                 ref_result_t * res)
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

	D3Q19ListAaPvInit_PushSoA(&ld, &kd, &p);
	/* Init leaves Iteration unset (src/BenchKernelD3Q19ListAaPvCommon.c:471-823);
	 * before the first step the PDFs are in the local (odd-parity) layout. */
	KDL(kd)->Iteration = -1;

	memset(res, 0, sizeof(*res));
	res->nCells = KDL(kd)->nCells;
	res->nConsecNodes = (int)KDLR(kd)->nConsecNodes;

	refh_init(&ld, kd, &cd, initMode, rho);
	if (iterations > 0) kd->Kernel(&ld, kd, &cd);
	refh_finish(&ld, kd, &cd, doVerify, dump, res);

	int nF = ld.nFluid;
	if (adj)    memcpy(adj, KDL(kd)->AdjList, sizeof(uint32_t) * (size_t)nF * N_D3Q19_IDX);
	if (grid)   memcpy(grid, KDL(kd)->Grid, sizeof(uint32_t) * (size_t)ld.nCells);
	if (coords) memcpy(coords, KDL(kd)->Coords, sizeof(uint32_t) * (size_t)nF * 3);

	D3Q19ListAaPvDeinit_PushSoA(&ld, &kd);
	MemFree((void **)&ld.Lattice);
	return 0;
}
