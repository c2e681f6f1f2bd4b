/*
 * lbm_host.h -- C ABI of liblbmhost.so: host-side counterparts of the
 * reference's shared helpers and geometry generator.
 *
 * These functions carry the reference's names and signatures so that a
 * driver written against the reference (src/Main.c) links unchanged.  They
 * reach the PDFs only through KernelData.GetNode/SetNode, exactly as the
 * reference's helpers do, so they work with any kernel of either library.
 * A program that already links the reference's Kernel.o and Geometry.o uses
 * those instead and needs only liblbmhip.so.
 */
#ifndef LBM_HOST_H
#define LBM_HOST_H

#include <stdint.h>

#include "lbm_kernel_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* D3Q19 tables (src/Kernel.h:163-168, src/Kernel.c:34-63). */
extern int D3Q19_X[D3Q19_NDIR];
extern int D3Q19_Y[D3Q19_NDIR];
extern int D3Q19_Z[D3Q19_NDIR];
extern int D3Q19_INV[D3Q19_NDIR];
extern const char * D3Q19_NAMES[D3Q19_NDIR];

/* src/Kernel.h:226 -- mean density over all fluid cells. */
PdfT KernelDensity(KernelData * kd, LatticeDesc * ld);
/* src/Kernel.h:228-229 -- density line on stdout, ux-progress.dat and
 * density-ux.dat in the working directory. */
void KernelStatistics(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration);
void KernelStatisticsAdv(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration, int forceOutput);
/* src/Kernel.h:232-233 -- equilibrium at rest, scaled by omega as the
 * reference writes it (sum f = omega * rho). */
void KernelSetInitialDensity(LatticeDesc * ld, KernelData * kd, CaseData * cd);
void KernelSetInitialVelocity(LatticeDesc * ld, KernelData * kd, CaseData * cd);
/* src/Kernel.h:235 -- L2 norm of the relative deviation of the x velocity
 * at x = nX/2 from analytic plane Poiseuille flow; writes flow-profile.dat. */
void KernelVerifiy(LatticeDesc * ld, KernelData * kd, CaseData * cd, PdfT * errorNorm);
/* src/Kernel.h:237 -- Luo body force in node view (host side).  The GPU
 * kernels apply it inside Kernel() themselves. */
void KernelAddBodyForce(KernelData * kd, LatticeDesc * ld, CaseData * cd);

/* src/Vtk.h:31 -- binary legacy-VTK snapshot vtk/file-<iteration>.vtk of
 * cell types, density and velocity (the reference's format byte for byte). */
void VtkWrite(LatticeDesc * ld, KernelData * kd, CaseData * cd, int iteration);

/* Geometry (src/Geometry.h:33-47) plus a seeded porous medium. */
typedef enum GEO_TYPES_ {
	GEO_TYPE_MIN = 0,
	GEO_TYPE_BOX = 0,
	GEO_TYPE_CHANNEL = 1,
	GEO_TYPE_PIPE = 2,
	GEO_TYPE_BLOCKS = 3,     /* typeDetails: int * block size */
	GEO_TYPE_FLUID = 4,
	GEO_TYPE_POROUS = 5,     /* typeDetails: GeoPorous *; no reference counterpart */
	GEO_TYPE_MAX = 5
} GEO_TYPES;

/* Porous medium: the channel geometry (fluid x faces, periodic in x, solid
 * y and z faces) filled with solid spheres of radius Radius whose centres
 * are drawn uniformly in [0,nX)x[0,nY)x[0,nZ) from a SplitMix64 stream
 * seeded with Seed; spheres wrap periodically in x.  Spheres are added
 * until nFluid / (nX nY nZ) <= Porosity.  A cell (x,y,z) belongs to a
 * sphere with centre c when (x-cx)^2 + (y-cy)^2 + (z-cz)^2 <= r^2 (x
 * distance taken to the nearest periodic image). */
typedef struct GeoPorous_ {
	double Radius;
	double Porosity;
	uint64_t Seed;
	int nSpheres;            /* out: spheres placed */
} GeoPorous;

void GeoCreateByType(GEO_TYPES type, void * typeDetails, int dims[3], int periodic[3], LatticeDesc * ld);
/* "box", "channel", "pipe", "blocks[-<n>]", "fluid" as the reference
 * (src/Geometry.c:37-88), and "porous[-<radius>[-<porosity>[-<seed>]]]". */
void GeoCreateByStr(const char * geometryType, int dims[3], int periodic[3], LatticeDesc * ld);
/* Frees ld->Lattice (the reference's driver does MemFree, src/Main.c:649). */
void GeoFree(LatticeDesc * ld);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HOST_H */
