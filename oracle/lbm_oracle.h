/*
 * lbm_oracle.h -- CPU restatement of the reference's D3Q19 TRT AA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it; the product (liblbmhip.so, liblbmhost.so, lbmbenchk-hip) never links,
 * loads or calls it.
 *
 * Every function restates one piece of the reference, cited file:line, in
 * plain C with the reference's scalar operation order and no FMA contraction
 * (built with -ffp-contract=off), so that its floating-point results are
 * bit-identical to the reference where the reference evaluates the same
 * expression tree.  The reference's list-aa-pv-soa kernel evaluates some
 * nodes with a different association tree (its AVX vector path,
 * src/BenchKernelD3Q19ListAaPv.c:242-378 and :610-754); orc_list_run can
 * reproduce that split exactly (vsize/nThreads arguments) or evaluate every
 * node with the scalar tree (vsize = 1).
 *
 * How each function is pinned (the tests under tests/ that do it):
 *   orc_pad_*            closed-form arithmetic + values measured on the
 *                        reference (test_oracle_basics.py) + live reference
 *                        (test_oracle_vs_ref.py)
 *   orc_geometry         cell counts in closed form + live reference
 *   orc_list_number,
 *   orc_list_adjacency   bijection/symmetry invariants, brute force on tiny
 *                        grids, golden fixtures and live reference
 *                        (bit-exact)
 *   orc_trt_collide      conservation of mass and momentum, rest state as a
 *                        fixed point, textbook TRT on random inputs
 *   orc_list_run,
 *   orc_aa_run           golden fixtures written by the reference build
 *                        (bit-exact for aa-soa and for list-aa-pv-soa with
 *                        the reference's vector split), mass conservation,
 *                        analytic Poiseuille profile
 *   orc_verify           the reference's own -verify deviation value
 *   orc_list_run_shifted,
 *   orc_aa_run_shifted   FP16 rounding against numpy's binary16, rest state
 *                        as a fixed point, tracking of the FP64 runs within
 *                        stated tolerances, run-split invariance, and the
 *                        two shifted runs bit-identical to each other where
 *                        their boundary handling coincides (no open faces)
 *                        (test_oracle_shifted.py)
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NDIR  19
#define ORC_NLINK 18

/* Stencil tables (src/Kernel.h:121-140, src/Kernel.c:34-63). */
extern const int    ORC_CX[ORC_NDIR];
extern const int    ORC_CY[ORC_NDIR];
extern const int    ORC_CZ[ORC_NDIR];
extern const int    ORC_INV[ORC_NDIR];
extern const double ORC_W[ORC_NDIR];   /* src/Kernel.c:769-772 */

/* ---- padding (src/Padding.c) ------------------------------------------ */
typedef struct {
	int nEntries;
	int modulus[16];
	int offset[16];
} orc_pad_t;

/* Parse "auto", "no"/"off" or "m+o(,m+o)*" (src/Padding.c:19-193).
 * Returns 0 on success, a negative code on a malformed string. */
int orc_pad_parse(const char * str, orc_pad_t * out);
/* Padded cell count (src/Padding.c:195-244); pad == NULL selects "auto". */
int orc_pad_cells(int nCells, int cellBytes, const orc_pad_t * pad);

/* ---- geometry (src/Geometry.c:90-281) --------------------------------- */
enum { ORC_GEO_BOX = 0, ORC_GEO_CHANNEL = 1, ORC_GEO_PIPE = 2, ORC_GEO_BLOCKS = 3, ORC_GEO_FLUID = 4 };
/* Fills mask[dims0*dims1*dims2] (x fastest).  periodicIn is the flag set
 * stored in LatticeDesc; periodicFaces receives the flags that decide whether
 * faces are fluid (the reference modifies its local copy after storing the
 * LatticeDesc flags, src/Geometry.c:115-139).  Returns nFluid. */
int orc_geometry(int type, int blockSize, const int dims[3], const int periodicIn[3],
                 int * mask, int periodicFaces[3]);

/* ---- list build (src/BenchKernelD3Q19ListAaPvCommon.c:587-793) -------- */
/* Node numbering over (optionally blocked) x,y,z loops with z fastest.
 * grid[L(x,y,z)] = node id or 0xffffffff for obstacles; coords[3n+k].
 * Returns nFluid. */
int  orc_list_number(const int * mask, const int dims[3], const int blk[3],
                     uint32_t * grid, uint32_t * coords);
/* Push adjacency, node-major adj[n*18 + d]. */
void orc_list_adjacency(const int * mask, const int dims[3], const int periodic[3],
                        const uint32_t * grid, const uint32_t * coords,
                        int nFluid, int nCells, uint32_t * adj);
/* Number of runs of nodes with consecutive adjacency, with forced breaks at
 * the reference's OpenMP node-range boundaries
 * (src/BenchKernelD3Q19ListAaPvCommon.c:326-438).  runLen may be NULL. */
int  orc_list_runs(const uint32_t * adj, int nFluid, int nThreads, uint32_t * runLen);

/* ---- collision (src/BenchKernelD3Q19ListAaPv.c:380-465) --------------- */
enum { ORC_TREE_SCALAR = 0, ORC_TREE_AVX = 1 };
/* Post-collision values post[d] for direction d (not yet swapped). */
void orc_trt_collide(const double pre[ORC_NDIR], double post[ORC_NDIR],
                     double omegaEven, double omegaOdd, int tree);
/* omegaOdd = 1 / (1/2 + Lambda / (1/omega - 1/2))
 * (src/BenchKernelD3Q19ListAaPv.c:177-178). */
double orc_omega_odd(double omega, double magic);

/* ---- list AA kernel (src/BenchKernelD3Q19ListAaPv.c:45-159) ----------- */
typedef struct {
	int dims[3];
	int nFluid;
	int nCells;
	const int * mask;
	const uint32_t * grid;
	const uint32_t * adj;     /* node-major, 18 per node */
	double * pdf;             /* 19 * nCells, SoA d*nCells + n */
	int iteration;            /* parity of the last completed step; -1 = none */
} orc_list_t;

/* Runs `iterations` steps rounded up to even (even, [force], odd, [force]).
 * forceOn mirrors a VERIFICATION build of the reference
 * (src/BenchKernelD3Q19ListAaPv.c:100-126).  vsize = 1 evaluates every node
 * with the scalar tree; vsize = 4 with nThreads = t reproduces the
 * reference's AVX split for -t t. */
void orc_list_run(orc_list_t * L, int iterations, double omega, double magic,
                  double xForce, int forceOn, int nThreads, int vsize);
void orc_list_get_node(const orc_list_t * L, int x, int y, int z, double out[ORC_NDIR]);
void orc_list_set_node(orc_list_t * L, int x, int y, int z, const double in[ORC_NDIR]);

/* ---- aa kernel (src/BenchKernelD3Q19AaCommon.c:309-637,
 *                 src/BenchKernelD3Q19Aa.c:40-548) ------------------------ */
typedef struct {
	int dims[3];
	int gdims[3];             /* dims + 2 (one ghost layer) */
	int periodic[3];
	const int * mask;
	double * pdf;             /* 19 * gX*gY*gZ, d*gX*gY*gZ + x*gY*gZ + y*gZ + z */
	int * bbSrc;
	int * bbDst;
	int nBB;
	int iteration;
} orc_aa_t;

/* Allocates pdf (zeroed, ghost cells included) and the fix-up lists. */
int  orc_aa_init(orc_aa_t * A, const int * mask, const int dims[3], const int periodic[3]);
void orc_aa_free(orc_aa_t * A);
void orc_aa_run(orc_aa_t * A, int iterations, double omega, double magic,
                double xForce, int forceOn);
void orc_aa_get_node(const orc_aa_t * A, int x, int y, int z, double out[ORC_NDIR]);
void orc_aa_set_node(orc_aa_t * A, int x, int y, int z, const double in[ORC_NDIR]);

/* ---- helpers of src/Kernel.c ------------------------------------------ */
/* Equilibrium the way KernelSetInitialDensity writes it (src/Kernel.c:258-340)
 * for a cell at x of nX; the caller stores it through its SetNode. */
void orc_initial_density_node(double rhoIn, double rhoOut, double omega, int x, int nX,
                              double out[ORC_NDIR]);
/* KernelSetInitialVelocity's node value (src/Kernel.c:344-438). */
void orc_initial_velocity_node(double omega, double out[ORC_NDIR]);
/* Luo body force term 3*w_d*c_x,d*F as the reference evaluates it
 * (src/Kernel.c:794-796). */
double orc_force_term(int d, double xForce);
/* Deviation norm of KernelVerifiy (src/Kernel.c:511-625) from the x velocity
 * slice ux[y*nZ + z] taken at x = nX/2 (solid cells 0), before the +F/2
 * correction, which this function adds for fluid cells as
 * src/Kernel.c:491-493 does. */
double orc_verify_norm(const double * uxSlice, const int * fluidSlice, int nY, int nZ,
                       double omega, double xForce);

/* Canonical node view: every fluid cell in mask order (x fastest), 19 values
 * each, taken through the parity-aware accessors. */
void orc_list_dump(const orc_list_t * L, double * out);

/* ---- shifted storage (no reference counterpart) ----------------------
 * Slots hold f - w_d in float; for FP16S the float holds (f - w_d) * 2^k
 * rounded to the nearest binary16 value.  Collision in float, force
 * increments added at the next load (at the store after the last odd
 * step).  Node views are valid between runs (complete step pairs). */
#define ORC_PREC_FP32S 1
#define ORC_PREC_FP16S 2
float orc_f16_round(float v);
void orc_list_run_shifted(const orc_list_t * L, float * store, int prec, int f16Shift, int iterations,
                          double omega, double magic, double xForce, int forceOn);
void orc_shifted_from_nodes(const orc_list_t * L, float * store, int prec, int f16Shift, const double * nodes);
void orc_shifted_to_nodes(const orc_list_t * L, const float * store, int prec, int f16Shift, double * nodes);
void orc_aa_dump(const orc_aa_t * A, double * out);
/* aa-soa over shifted storage laid out like orc_aa_t.pdf (float per slot,
 * every slot -- ghost cells included -- starting at f = 0).  Returns -1 when
 * the scratch copy of the last odd step cannot be allocated. */
int  orc_aa_run_shifted(const orc_aa_t * A, float * store, int prec, int f16Shift, int iterations,
                        double omega, double magic, double xForce, int forceOn);
void orc_aa_shifted_from_nodes(const orc_aa_t * A, float * store, int prec, int f16Shift, const double * nodes);
void orc_aa_shifted_to_nodes(const orc_aa_t * A, const float * store, int prec, int f16Shift, double * nodes);

#ifdef __cplusplus
}
#endif

#endif /* LBM_ORACLE_H */
