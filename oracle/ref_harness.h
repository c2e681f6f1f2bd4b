/*
 * ref_harness.h -- entry points of oracle/_ref/libref_*.so.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
 * sources (Kernel.c, Geometry.c, Padding.c, Memory.c, Lattice.c and the
 * list-aa-pv-soa / aa-soa kernel files, unmodified, read from where they lie)
 * together with the glue in ref_harness_*.c.  The glue drives the reference
 * the way its driver does (src/Main.c:476-592): build a LatticeDesc, call the
 * kernel's Init, initialise through SetNode or the reference helpers, call
 * kd->Kernel, then read every fluid node back through kd->GetNode.
 *
 * Two variants are built: libref_verif.so (VERIFICATION defined: body force
 * after every step, src/BenchKernelD3Q19ListAaPv.c:100-126) and
 * libref_bench.so (the reference's default benchmark build, no forcing).
 */
#ifndef REF_HARNESS_H
#define REF_HARNESS_H

#include <stdint.h>

/* init modes */
enum {
	REF_INIT_NONE = 0,       /* PDFs as Init leaves them (all 0.0) */
	REF_INIT_VERIFY = 1,     /* KernelSetInitialDensity + KernelSetInitialVelocity (src/Main.c:562-567) */
	REF_INIT_WEIGHTS = 2     /* SetNode(f_d = w_d * rho) on every fluid cell */
};

typedef struct {
	int nFluid;
	int nCells;              /* list: padded node count; aa: gX*gY*gZ */
	int nBounceBack;         /* aa only */
	int nConsecNodes;        /* list only: runs reported by the reference */
	double duration;         /* kd->Duration, seconds */
	double verifyNorm;       /* KernelVerifiy result, or -1 when not requested */
	double density;          /* KernelDensity after the run */
} ref_result_t;

/* Runs the reference list-aa-pv-soa kernel.  mask: int32 x-fastest.
 * kernelArgs: the words after "--" (may be NULL).  Outputs may be NULL:
 *   dump   nFluid*19 doubles, fluid cells in mask order via GetNode
 *   adj    nFluid*18 uint32, node-major (KernelDataList.AdjList)
 *   grid   nX*nY*nZ uint32 (obstacles undefined), coords 3*nFluid uint32 */
int ref_list_run(const int * mask, const int dims[3], const int periodic[3],
                 int nThreads, int nKernelArgs, char ** kernelArgs,
                 double omega, double xForce, int iterations, int initMode, double rho,
                 int doVerify,
                 double * dump, uint32_t * adj, uint32_t * grid, uint32_t * coords,
                 ref_result_t * res);

/* Runs the reference aa-soa kernel; same conventions. */
int ref_aa_run(const int * mask, const int dims[3], const int periodic[3],
               int nThreads, int nKernelArgs, char ** kernelArgs,
               double omega, double xForce, int iterations, int initMode, double rho,
               int doVerify, double * dump, ref_result_t * res);

/* GeoCreateByStr on the reference; mask receives nX*nY*nZ ints, ldPeriodic
 * the flags stored in the LatticeDesc.  Returns nFluid. */
int ref_geometry(const char * type, const int dims[3], const int periodic[3],
                 int * mask, int ldPeriodic[3]);

/* CaseData.VtkOutput / VtkModulus / StatisticsModulus of the following runs
 * (VTK and statistics files appear only with libref_io.so, in the working
 * directory). */
void ref_set_output(int vtkOutput, int vtkModulus, int statisticsModulus);

/* PadCells(nCells, sizeof(double), PadInfoFromStr(str)); str NULL = default. */
int ref_pad_cells(int nCells, const char * str);

#endif /* REF_HARNESS_H */
