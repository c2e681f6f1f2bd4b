/*
 * lbm_kernel_abi.h -- the reference's kernel-plugin interface, re-declared.
 *
 * The reference selects one of its kernels at run time through a table of
 * KernelFunctions entries (src/KernelFunctions.h:49-54); each entry's Init
 * fills a KernelData (src/Kernel.h:187-211) whose function pointers the
 * driver and the shared helpers then call.  The types below are declared with
 * the same field order, field types and sizes as
 *
 *   LatticeDesc      src/Lattice.h:43-55
 *   CaseData         src/Kernel.h:172-184
 *   KernelData       src/Kernel.h:187-211
 *   Parameters       src/Kernel.h:217-222
 *   KernelFunctions  src/KernelFunctions.h:49-54
 *
 * so that an object compiled against the reference headers (its Main.c,
 * Kernel.c helpers, Vtk.c writer) and an object compiled against this header
 * (the HIP kernels in liblbmhip.so) exchange these structs unchanged.  PdfT
 * is double: the reference's PRECISION_DP build (src/Kernel.h:73-74), which is
 * the precision of the path accelerated here.
 *
 * Plain C, no HIP or torch types.
 */
#ifndef LBM_KERNEL_ABI_H
#define LBM_KERNEL_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double PdfT;
typedef int    LatticeT;

/* Cell types of the geometry mask (src/Lattice.h:35-40). */
enum {
	LBM_CELL_OBSTACLE = 0,
	LBM_CELL_FLUID    = 1,
	LBM_CELL_INLET    = 2,
	LBM_CELL_OUTLET   = 4
};

/* D3Q19 direction numbering (src/Kernel.h:87-111): the rest direction C is
 * the last one, the 18 moving directions come first. */
enum {
	D3Q19_DIR_N = 0,  D3Q19_DIR_S = 1,  D3Q19_DIR_E = 2,  D3Q19_DIR_W = 3,
	D3Q19_DIR_NE = 4, D3Q19_DIR_SE = 5, D3Q19_DIR_NW = 6, D3Q19_DIR_SW = 7,
	D3Q19_DIR_T = 8,  D3Q19_DIR_TN = 9, D3Q19_DIR_TE = 10, D3Q19_DIR_TW = 11,
	D3Q19_DIR_TS = 12,
	D3Q19_DIR_B = 13, D3Q19_DIR_BS = 14, D3Q19_DIR_BN = 15, D3Q19_DIR_BW = 16,
	D3Q19_DIR_BE = 17,
	D3Q19_DIR_C = 18,
	D3Q19_NDIR = 19,
	D3Q19_NLINK = 18      /* moving directions = adjacency entries per node */
};

/* Geometry description: int32 cell mask, x fastest (src/Lattice.h:43-55). */
typedef struct LatticeDesc_ {
	int Dims[3];
	LatticeT * Lattice;
	int nCells;
	int nFluid;
	int nObst;
	int nInlet;
	int nOutlet;
	int PeriodicX;
	int PeriodicY;
	int PeriodicZ;
	const char * Name;
} LatticeDesc;

/* Flow parameters handed to KernelData.Kernel (src/Kernel.h:172-184). */
typedef struct CaseData_ {
	PdfT Omega;
	PdfT RhoIn;
	PdfT RhoOut;
	PdfT Ux;
	PdfT Uy;
	PdfT Uz;
	PdfT XForce;
	int MaxIterations;
	int VtkOutput;
	int VtkModulus;
	int StatisticsModulus;
} CaseData;

/* Per-kernel state; derived kernel structs embed it as their first member
 * (src/Kernel.h:187-211). */
typedef struct KernelData_ {
	PdfT * Pdfs[2];
	PdfT * SrcPdfs;
	PdfT * DstPdfs;
	PdfT * PdfsActive;
	int Dims[3];
	int GlobalDims[3];
	int Offsets[3];
	int * ObstIndices;
	int nObstIndices;
	int * BounceBackPdfsSrc;
	int * BounceBackPdfsDst;
	int nBounceBackPdfs;
	double LoopBalance;
	double Duration;

	void (* BoundaryConditionsGetPdf)(struct KernelData_ * kd, int x, int y, int z, int dir, PdfT * pdf);
	void (* BoundaryConditionsSetPdf)(struct KernelData_ * kd, int x, int y, int z, int dir, PdfT pdf);

	void (* GetNode)(struct KernelData_ * kd, int x, int y, int z, PdfT * pdfs);
	void (* SetNode)(struct KernelData_ * kd, int x, int y, int z, PdfT * pdfs);

	void (* Kernel)(LatticeDesc * ld, struct KernelData_ * kd, CaseData * cd);
} KernelData;

/* Command line split at "--": the words after it go to the kernel's Init
 * (src/Kernel.h:217-222, filled at src/Main.c:380-385). */
typedef struct Parameters_ {
	int nArgs;
	char ** Args;
	int nKernelArgs;
	char ** KernelArgs;
} Parameters;

/* One entry of a kernel table (src/KernelFunctions.h:49-54). */
typedef struct KernelFunctions_ {
	char Name[128];
	void (* Init)(LatticeDesc * ld, KernelData ** kernelData, Parameters * params);
	void (* Deinit)(LatticeDesc * ld, KernelData ** kernelData);
} KernelFunctions;

#ifdef __cplusplus
}
#endif

#endif /* LBM_KERNEL_ABI_H */
