/*
 * lbm_hip.h -- C ABI of liblbmhip.so, the MI355X (gfx950) implementation of
 * the reference's D3Q19 TRT AA-pattern kernels.
 *
 * Drop-in entry points.  Each replaces one kernel-table function of the
 * reference and has the same signature and argument meaning; the KernelData
 * it returns carries Kernel/GetNode/SetNode/BoundaryConditions{Get,Set}Pdf
 * function pointers with the reference's semantics (src/Kernel.h:187-211).
 * Errors behave as in the reference: "ERROR: ..." on stdout and exit(1),
 * unless LbmHipSetErrorMode(LBM_HIP_ERRORS_RETURN) was called, in which case
 * Init leaves *kernelData NULL and LbmHipLastError() holds the message.
 *
 * Kernel arguments (the words after "--", Parameters.KernelArgs):
 *   -blk <n> | -blk-x/-blk-y/-blk-z <n>   node-numbering blocks (list only,
 *                                         src/BenchKernelD3Q19ListAaPvCommon.c:236-301)
 *   -pad auto|no|<m>+<o>(,<m>+<o>)*       SoA node padding (list only,
 *                                         src/Padding.c:19-193)
 *   -force on|off     apply the Luo body force after every step, as the
 *                     reference's VERIFICATION build does (default on)
 *   -lambda <v>       TRT magic parameter (default 1/12, the value the
 *                     reference hard-codes, src/BenchKernelD3Q19ListAaPv.c:177)
 *   -bs <n>           threads per block of the step kernels (64..512, step 64;
 *                     list-aa-pv-soa -- aa-soa's blocks follow its z rows)
 *   -adj rle|full     list only: the odd step reads the adjacency run-compressed
 *                     (default) or in full (18 entries per node)
 *   -nt on|off        non-temporal loads/stores in the even step (default:
 *                     on when what a step touches -- PDFs, and for the list
 *                     kernel its adjacency -- exceeds 256 MiB, the device's
 *                     last-level cache; off for cache-sized lattices, which
 *                     stream faster with plain accesses)
 *   -even-cap on|off|force
 *                     even step in 512-thread blocks, one block per compute
 *                     unit (default on for lattices of at least 16 such
 *                     blocks per compute unit, except list-aa-pv-soa with
 *                     four FP16 nodes per thread and aa-soa with shifted
 *                     storage over 256 MiB; force: at any size; -bs then
 *                     applies to the odd step)
 *   -prec fp64|fp32s|fp16s
 *                     PDF storage (both kernels).  fp64 (default) is the
 *                     reference's PdfT = double.  fp32s/fp16s store the
 *                     deviation f - w_d (FP16: times 2^k, see -fp16-scale)
 *                     and collide in FP32; no counterpart in the reference,
 *                     whose single precision stores f in float
 *                     (src/Kernel.h:75-76).  GetNode/SetNode still exchange
 *                     double node views; KernelData.Pdfs[0] then points to
 *                     the storage type.  The body-force increment is added
 *                     in FP32 at the next step's load (after the last step of
 *                     a Kernel() call, at the store).  aa-soa's ghost layer
 *                     starts at f = 0 in every storage type.
 *   -fp16-scale <2^k> power-of-two factor of the FP16 deviations
 *                     (default 4096)
 *   -statistics       call the program's KernelStatistics during the run as
 *                     the reference's STATISTICS build does; with
 *                     CaseData.VtkOutput set the program's VtkWrite is called
 *                     at the reference's points as well, with or without
 *                     -statistics (src/BenchKernelD3Q19ListAaPv.c:67-156,
 *                     src/BenchKernelD3Q19Aa.c:102-112, 296-332, 513-547)
 *   -shadow eager|lazy list only: allocate the host copy at Init (default) or
 *                     at the first LbmHipSyncShadow / download
 *   -odd-order band|group|plain
 *                     list only: order of the odd step's blocks over the
 *                     device's 8 XCDs -- band (default): each XCD takes the
 *                     blocks of one eighth of the y range; group: 8
 *                     consecutive blocks per XCD in turn; plain: launch order
 *   -even-order up|down
 *                     list only: even-step blocks in ascending (default) or
 *                     descending node order (starting on the planes the odd
 *                     step ended on)
 *   -graph on|off     capture the step launches of a Kernel() call as a HIP
 *                     graph and replay it while steps, omega and force stay
 *                     the same (default off; not with -launch-events, VTK
 *                     output or -statistics, which run the plain loop)
 *   -quiet            no informational lines at Init
 *   -launch-events on|off
 *                     time every even and odd launch with HIP events
 *                     (LbmHipTimings even/odd averages; default off)
 *   -vec auto|1       fp32s/fp16s: the even step takes 2 (FP32) or 4 (FP16)
 *                     nodes (aa-soa: consecutive z cells) per thread with
 *                     8-byte accesses (auto, default; the list kernel only
 *                     when the padded node count allows), or one per
 *                     thread (1)
 *   -odd-pairs auto|on|off
 *                     aa-soa at fp32s with -vec auto: the odd step takes
 *                     pairs of z cells, one access per direction for a pair
 *                     of interior fluid cells (auto, default: on lattices
 *                     over 256 MiB, where it measured faster; one cell per
 *                     thread on smaller ones)
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stdint.h>

#include "lbm_kernel_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-in kernel-table functions ------------------------------------ */

/* Replaces D3Q19ListAaPvInit_PushSoA (src/BenchKernelD3Q19ListAaPv.h:32,
 * table entry "list-aa-pv-soa" src/KernelFunctions.h:58-62).  Builds the
 * fluid-node list and its 18-entry push adjacency on the GPU, bit-identical
 * to the reference's list build, and allocates the device lattice. */
void D3Q19ListAaPvInit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params);

/* Replaces D3Q19ListAaPvDeinit_PushSoA (src/BenchKernelD3Q19ListAaPv.h:34). */
void D3Q19ListAaPvDeinit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData);

/* Replaces D3Q19AaInit_AaSoA (src/BenchKernelD3Q19Aa.h:33, table entry
 * "aa-soa" src/KernelFunctions.h:171-175). */
void D3Q19AaInit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params);

/* Replaces D3Q19AaDeinit_AaSoA (src/BenchKernelD3Q19Aa.h:36). */
void D3Q19AaDeinit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData);

/* Kernel table of this library, laid out like g_kernels[]
 * (src/KernelFunctions.h:56-188): "list-aa-pv-soa" and "aa-soa". */
extern KernelFunctions g_lbmHipKernels[];
extern const int g_lbmHipKernelCount;

/* ---- extensions (no counterpart in the reference) ----------------------- */

enum { LBM_HIP_ERRORS_EXIT = 0, LBM_HIP_ERRORS_RETURN = 1 };
/* Selects the error behaviour of every entry point (process-wide). */
void LbmHipSetErrorMode(int mode);
/* Message of the last error, "" if none.  Cleared by LbmHipClearError. */
const char * LbmHipLastError(void);
void LbmHipClearError(void);

/* Per-call timing of the last Kernel() call, from HIP events recorded on
 * the kernel's stream around every step launch. */
typedef struct {
	int    steps;          /* steps run (MaxIterations rounded up to even) */
	double loopMs;         /* first launch start to last launch end */
	double evenMsAvg;      /* average duration of one even-step launch */
	double oddMsAvg;       /* average duration of one odd-step launch */
	double evenMsMin, oddMsMin;
	double uploadMs;       /* host shadow -> device before the loop (0 if clean) */
	double downloadMs;     /* device -> host shadow after the loop */
} LbmHipTimings;
int LbmHipGetTimings(const KernelData * kd, LbmHipTimings * out);

/* Enables/disables the per-launch events (off by default, or as set by
 * -launch-events).  With them off only loopMs is measured; with them on the
 * even/odd averages are too, at the cost of an event record between
 * launches (about 4 us per step: 25 % of a step at 64^3, 0.3 % at 256^3). */
int LbmHipSetLaunchEvents(KernelData * kd, int on);
/* Enables/disables the device -> host shadow copy at the end of Kernel()
 * (on by default, needed by GetNode).  With it off, GetNode reads the state
 * before the last Kernel() call until LbmHipSyncShadow is called. */
int LbmHipSetShadowDownload(KernelData * kd, int on);
int LbmHipSyncShadow(KernelData * kd);

/* Sizes of the list kernel: fluid nodes, padded node stride nCells,
 * adjacency row stride on the device, and the number of runs of nodes
 * with consecutive adjacency (src/BenchKernelD3Q19ListAaPvCommon.c:326-469,
 * without thread breaks). */
typedef struct {
	int nFluid;
	int nCells;
	int adjStride;
	int nRuns;
	int dims[3];
	int periodic[3];
	double bytesPerFlupEven;
	double bytesPerFlupOdd;    /* algorithmic bytes per node of the odd step as configured */
	int nRunsWave;             /* runs when cut at every 64-node wavefront (run-compressed odd step) */
	int adjRle;                /* 1: odd step reads the run-compressed adjacency (-adj rle) */
	int prec;                  /* PDF storage (-prec): 0 fp64, 1 fp32s, 2 fp16s */
} LbmHipListInfo;
int LbmHipListGetInfo(const KernelData * kd, LbmHipListInfo * out);

/* FP64 sums over the fluid nodes of the current lattice, computed on the
 * device: mass = sum_n sum_d f_d and momentumX = sum_n sum_d c_x,d f_d
 * (node views; the counterparts of the per-node loops of KernelDensity and
 * KernelStatisticsAdv, src/Kernel.c:209-254, 633-755).  Valid between
 * Kernel() calls; pending SetNode changes are uploaded first.
 * LbmHipMoments takes either kernel (list-aa-pv-soa or aa-soa);
 * LbmHipListMoments only the list kernel. */
int LbmHipMoments(KernelData * kd, double * mass, double * momentumX);
int LbmHipListMoments(KernelData * kd, double * mass, double * momentumX);

/* Runs the case in Kernel() calls of checkEvery steps (rounded up to even)
 * until the relative change of the x momentum between two checks is <= tol
 * or maxSteps steps have run.  The host shadow is refreshed once at the end
 * (when shadow download is on); kd->Duration is the summed step time.
 * No counterpart in the reference, which always runs MaxIterations steps
 * (src/Main.c:574).  LbmHipRunSteady takes either kernel,
 * LbmHipListRunSteady only the list kernel. */
typedef struct {
	int steps;          /* steps run */
	int converged;      /* 1 when the tolerance was met */
	double relChange;   /* last relative change of momentumX (-1 before two checks) */
	double mass;        /* LbmHipMoments after the last chunk */
	double momentumX;
	double seconds;     /* summed step-loop time */
} LbmHipSteady;
int LbmHipRunSteady(KernelData * kd, CaseData * cd, int checkEvery, double tol, int maxSteps, LbmHipSteady * out);
int LbmHipListRunSteady(KernelData * kd, CaseData * cd, int checkEvery, double tol, int maxSteps, LbmHipSteady * out);

/* ---- batch: many list-aa-pv-soa lattices stepped together --------------
 * No counterpart in the reference (one lattice per process).  Every domain
 * is an ordinary list-aa-pv-soa instance created with the batch's kernel
 * arguments (LbmHipBatchKernelData gives its KernelData: GetNode, SetNode,
 * LbmHipListMoments all apply), initialised to f = w (rho 1, u 0).  A step
 * interval advances every active domain by the same number of steps in one
 * even and one odd launch per step over a list of node chunks of all of
 * them (replayed from a HIP graph) and then checks each domain's x momentum
 * in FP64; a domain leaves the active set when the relative change between
 * two checks is <= tol or its maxSteps (checked at interval ends) is
 * reached.  Each domain's result is bit-identical to the same lattice run
 * alone through Kernel() for the same number of steps (FP64 and fp32s;
 * for fp16s, to Kernel() calls of one interval each).  After stepping, the
 * host copies are stale until LbmHipBatchSync.  Not thread safe. */
typedef struct LbmHipBatch LbmHipBatch;
typedef struct {
	int active;         /* 1 while the domain is stepped */
	int converged;      /* 1 when it met the tolerance */
	int steps;          /* steps run */
	int nFluid;
	double relChange;   /* last relative change of momentumX (-1 before two checks) */
	double mass;        /* FP64 sums after the last interval */
	double momentumX;
	double seconds;     /* GPU time of the intervals it took part in */
} LbmHipBatchStatus;
typedef struct {
	int evenBlock;      /* threads per block of the even step (64..512) */
	int evenCap;        /* 1: one even-step block per compute unit (LDS reservation) */
	int vec;            /* nodes per thread of the even step: 1, 2 (fp32s), 4 (fp16s) */
	int oddBlock;       /* threads per block of the odd step (64..512) */
	double evenMs, oddMs;   /* per launch, as measured by LbmHipBatchAutotune */
	int groupMB;        /* > 0: domains stepped in groups of about this many MiB,
	                       each group through a whole interval before the next
	                       (its lattices stay in the last-level cache; default
	                       256); 0, or all domains within one group: every
	                       launch covers all domains.  Either way an
	                       interval is replayed from captured graphs, one
	                       per group */
	double pairMs;      /* per step pair with the chosen groups (autotune) */
	int nontemporal;    /* even-step loads/stores: 1 non-temporal, 0 plain,
	                       -1 (default, unless the batch was created with -nt
	                       on|off) per group: non-temporal only when the
	                       group's fluid PDFs and adjacency exceed the 256 MiB
	                       last-level cache */
	int oddGroup;       /* odd step: consecutive chunks each XCD takes per turn
	                       (a power of two up to 256; 0: 8) -- neighbouring
	                       chunks of one domain then share an L2 */
} LbmHipBatchTuning;
/* params: kernel arguments shared by all domains (-prec, -lambda, -adj, ...). */
LbmHipBatch * LbmHipBatchCreate(const Parameters * params);
void LbmHipBatchDestroy(LbmHipBatch * b);
/* Adds a domain; returns its id (>= 0) or -1.  maxSteps <= 0: no limit. */
int LbmHipBatchAdd(LbmHipBatch * b, LatticeDesc * ld, double omega, double xForce, int maxSteps);
/* Frees a domain (active or not); its id may be reused by a later Add. */
int LbmHipBatchRemove(LbmHipBatch * b, int id);
KernelData * LbmHipBatchKernelData(LbmHipBatch * b, int id);
int LbmHipBatchStatusOf(LbmHipBatch * b, int id, LbmHipBatchStatus * out);
/* One interval of `steps` (rounded up to even) for all active domains;
 * *nActive receives the number still active afterwards. */
int LbmHipBatchStep(LbmHipBatch * b, int steps, double tol, int * nActive);
/* The two halves of LbmHipBatchStep: Launch starts the interval on the
 * batch's stream and returns; Wait completes it (convergence test, status).
 * Between them the batch accepts no other call; other batches and host work
 * may proceed, so intervals of several batches overlap on the device. */
int LbmHipBatchLaunch(LbmHipBatch * b, int steps);
int LbmHipBatchWait(LbmHipBatch * b, double tol, int * nActive);
/* Intervals of checkEvery steps until no domain is active. */
int LbmHipBatchRun(LbmHipBatch * b, int checkEvery, double tol);
/* Device -> host copy of one domain (id >= 0) or of all (id < 0). */
int LbmHipBatchSync(LbmHipBatch * b, int id);
int LbmHipBatchGetTuning(LbmHipBatch * b, LbmHipBatchTuning * out);
int LbmHipBatchSetTuning(LbmHipBatch * b, const LbmHipBatchTuning * t);
/* Times `pairs` launches of every candidate shape on the active domains and
 * keeps the fastest, then `pairs` step pairs for groups of 0 (none), 64,
 * 128, 192, 256, 384, 512, 768 and 1024 MiB with that shape, each with plain
 * and with non-temporal even-step accesses, then the shapes again under the
 * chosen groups; the lattices are restored afterwards.  Odd-step shapes
 * include oddGroup 8, 32 and 128. */
int LbmHipBatchAutotune(LbmHipBatch * b, int pairs, LbmHipBatchTuning * out);
/* Summed GPU time of all intervals and fluid-node updates done. */
int LbmHipBatchStats(LbmHipBatch * b, double * gpuSeconds, double * flups);
/* Graph bookkeeping of the intervals: the launch ranges of the last interval
 * (1, or the number of groups), graphs captured and graph replays so far.
 * A range keeps its graph while the step count, the launch shapes and its
 * device tables stay; domains leaving or joining only rewrite the tables. */
int LbmHipBatchGraphStats(LbmHipBatch * b, int * slots, long long * captures, long long * replays);

/* Upper bound of the device memory D3Q19ListAaPvInit_PushSoAHip takes for
 * the lattice ld with these kernel arguments (padding, storage type, ...),
 * computed on the host without touching the device: every allocation
 * Init makes, its temporary buffers included, the run-compressed adjacency
 * at its largest, each rounded up to 2 MiB.  For admission against a
 * memory budget (the serving loop). */
int LbmHipListDeviceBytes(const LatticeDesc * ld, const Parameters * params, uint64_t * bytes);

/* Copies the device list build to caller buffers, in the reference's
 * layouts (any pointer may be NULL):
 *   adj    nFluid*18 uint32, node-major (KernelDataList.AdjList)
 *   grid   nX*nY*nZ uint32, node id per cell, 0xffffffff for obstacles
 *   coords 3*nFluid uint32 */
int LbmHipListCopyBuild(const KernelData * kd, uint32_t * adj, uint32_t * grid, uint32_t * coords);

/* The list build of D3Q19ListAaPvInit_PushSoAHip (node numbering, grid,
 * coords, 18-entry push adjacency; bit-identical to it and to the
 * reference's, src/BenchKernelD3Q19ListAaPvCommon.c:627-793) into device
 * buffers the caller owns, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops with preallocated pools, harnesses
 * that forbid hipMalloc).
 *   dMask     device, dims[0]*dims[1]*dims[2] int32 cell types in the
 *             LatticeDesc.Lattice order (x fastest)
 *   periodic  PeriodicX/Y/Z; blk node-numbering blocks (NULL or 0: whole
 *             extent, as -blk)
 *   nCells    PDF node stride the adjacency values encode (d * nCells + m)
 *   adjStride row stride of dAdj, laid out [18][adjStride]
 *   capacity  nodes dCoords and dAdj have room for
 * Outputs (device): dGrid[cells] node id per cell (0xffffffff for
 * obstacles), dCoords[3 * nFluid], dAdj rows 0..nFluid-1 (entries past
 * nFluid are not written), *nFluid (host).  Fails, writing nothing, when the
 * lattice has more than 67,108,608 cells (the scratch limit) or more fluid
 * cells than capacity, nCells or adjStride.  Runs on the null stream and
 * returns when done.  Its 1 MiB scratch is a static device array shared by
 * every call on a GPU: calls that overlap on one GPU race. */
int LbmHipListBuildNoAlloc(const int * dMask, const int dims[3], const int periodic[3], const int blk[3],
                           int nCells, int adjStride, int capacity, uint32_t * dGrid, uint32_t * dCoords,
                           uint32_t * dAdj, int * nFluid);

/* Node views of every non-solid cell of ld, in mask order (x fastest),
 * 19 values each, through kd->GetNode / kd->SetNode. */
int LbmHipDumpNodes(KernelData * kd, const LatticeDesc * ld, PdfT * out);
int LbmHipLoadNodes(KernelData * kd, const LatticeDesc * ld, const PdfT * in);

/* Streaming-bandwidth yardsticks on this device, the way the reference
 * rates its kernels against copy and "update-19" bandwidth
 * (doc/main.rst:508-511): a copy of nDoubles doubles into a second array,
 * and an in-place read-modify-write of 19 arrays of nNodes doubles each.
 * Results in GB/s (1e9 B/s), averaged over reps launches after warm-up. */
typedef struct {
	double copyGBs;
	double update19GBs;
	double copyBytes;        /* bytes read + written by one copy launch */
	double update19Bytes;    /* bytes read + written by one update-19 launch */
} LbmHipStreamResult;
int LbmHipStreamBench(uint64_t nDoubles, uint32_t nNodes, int reps, LbmHipStreamResult * out);

/* Name and compute-unit count of the device the library runs on. */
int LbmHipDeviceInfo(char * name, int nameLen, int * computeUnits, double * memClockGHz);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
