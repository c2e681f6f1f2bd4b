// lbm_runtime.h -- host-side plumbing shared by the list and aa kernels of
// liblbmhip.so: error handling, kernel-argument parsing, SoA padding, the
// collision constants, the Luo force table, and the common per-kernel state.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stddef.h>

#include "lbm_hip.h"

namespace lbm {

// ---------------------------------------------------------------- errors --
// Records the message; in LBM_HIP_ERRORS_EXIT mode prints "ERROR: <msg>"
// and exits with status 1 like the reference's error paths
// (src/BenchKernelD3Q19ListAaPvCommon.c:246-317, src/Memory.c:63-66).
// Always returns false so callers can `return fail(...)`.
bool fail(const char * fmt, ...) __attribute__((format(printf, 1, 2)));
bool hip_ok(hipError_t e, const char * what);

// --------------------------------------------------------------- padding --
// Padding specification as the reference's PadInfo (src/Padding.h:11-17).
struct PadSpec {
	int n = 0;
	int mod[16] = {0};
	int off[16] = {0};
};
// "auto" | "no" | "off" | "<m>+<o>(,<m>+<o>)*" (src/Padding.c:19-193).
bool pad_parse(const char * s, PadSpec * out);
// Node count padded as src/Padding.c:195-244 does; ps == nullptr = "auto".
long pad_cells(long nCells, int cellBytes, const PadSpec * ps);

// ----------------------------------------------------------- kernel args --
struct KernelArgs {
	int blk[3] = {0, 0, 0};
	bool havePad = false;
	PadSpec pad;
	int forceOn = 1;
	double magic = 1.0 / 12.0;
	int blockSize = 256;
	int adjRle = 1;        // list: odd step reads run-compressed adjacency
	int nontemporal = -1;  // non-temporal loads/stores in the even step: 1 on, 0 off, -1 the kernel's default
	int evenCap = 1;       // even step: 512-thread blocks, one block per CU (2: also on small lattices)
	int evenCapGiven = 0;  // -even-cap was on the command line (else the kernel's default applies)
	int prec = 0;          // list: PDF storage, see enum Prec
	int f16Shift = 12;     // FP16 storage holds (f - w) * 2^f16Shift
	int evenVec = 1;       // narrow storage: several nodes per thread in the even step
	int oddPairs = -1;     // aa-soa, FP32 storage: odd step over pairs of z cells (1 on, 0 off,
	                       // -1 on lattices over 256 MiB, where it measured faster)
	int quiet = 0;         // no informational lines at Init
	int lazyShadow = 0;    // list: host copy allocated at the first download, not at Init
	int statistics = 0;    // list: KernelStatistics during the run (the reference's STATISTICS build)
	int oddOrder = 2;      // list: odd-step block order, 0 plain, 1 XCD-grouped, 2 y bands per XCD
	int evenDown = 0;      // list: even step in descending block order (starts where the odd step ended)
	int launchEvents = 0;  // per-launch timing events (even/odd averages)
	int graph = 0;         // replay the step loop of a Kernel() call from a captured HIP graph
};

// PDF storage of the list kernel.
enum Prec { PREC_FP64 = 0, PREC_FP32S = 1, PREC_FP16S = 2 };

// MI355X's last-level (memory-side) cache.  A lattice whose PDFs fit stays
// cached from step to step, and the streaming even step then runs faster with
// plain than with non-temporal accesses (which would push it out); above this
// size non-temporal accesses are faster.  Both kernels' -nt default.
constexpr size_t kLastLevelCacheBytes = size_t(256) << 20;

// IEEE binary16 conversions on the host, round to nearest even (the
// rounding of the device's v_cvt_f16_f32).
uint16_t f32_to_f16(float f);
float f16_to_f32(uint16_t h);
// Parses Parameters.KernelArgs.  listKernel enables -pad (the reference
// accepts it only for the list kernel, src/BenchKernelD3Q19ListAaPvCommon.c:302-308).
bool parse_kernel_args(const Parameters * p, KernelArgs * a, bool listKernel);

// -------------------------------------------------------------- physics --
// TRT relaxation constants, each evaluated on the host with the operation
// order of src/BenchKernelD3Q19ListAaPv.c:174-189 so that the device sees
// the reference's exact double values.
struct TrtConsts {
	double oe, oo;        // omegaEven, omegaOdd
	double w0, w1, w2;
	double w1x3, w1nh;    // w_1 * 3, w_1 * 9 / 2
	double w2x3, w2nh;    // w_2 * 3, w_2 * 9 / 2
};
TrtConsts trt_consts(double omega, double magic);

// Luo body-force increments 3 * w_d * c_x,d * F per direction, evaluated
// left to right as src/Kernel.c:795 does.
struct ForceTerms {
	double f[19];
};
ForceTerms force_terms(double xForce);

extern const int kCx[19], kCy[19], kCz[19], kInv[19];

// ------------------------------------------------------- common state --
// Embedded (after KernelData) in the list and aa kernel structs.
struct Common {
	int device = 0;
	int ldsPerCU = 0;           // bytes of LDS per compute unit
	int ldsPerBlock = 0;        // largest dynamic LDS a block may request
	int nCU = 0;                // compute units of the device
	hipStream_t stream = nullptr;
	KernelArgs args;
	int iteration = -1;         // parity of the last completed step
	int shadowDirty = 0;        // SetNode touched the host shadow since the last upload
	int shadowStale = 0;        // the device lattice is newer than the host shadow
	int launchEvents = 0;       // from -launch-events; LbmHipSetLaunchEvents
	int shadowDownload = 1;
	hipEvent_t * ev = nullptr;  // 2 per launch
	int nEv = 0;
	hipEvent_t evLoop[2] = {nullptr, nullptr};
	LbmHipTimings t = {};
	bool (*download)(KernelData * kd) = nullptr;   // device lattice -> host shadow
	// the step loop of the last Kernel() call, captured (-graph on), and the
	// values it was captured for
	hipGraph_t graph = nullptr;
	hipGraphExec_t graphExec = nullptr;
	int graphSteps = 0;
	double graphOmega = 0.0, graphForce = 0.0;
};

// SetNode may change the host shadow only while it holds the device's
// state; otherwise it reports an error and returns false.
bool shadow_writable(const Common * c);

// Common state of a KernelData created by this library, or nullptr.
Common * common_of(KernelData * kd);
Common * list_common(KernelData * kd);
Common * aa_common(KernelData * kd);

bool common_init(Common * c, const Parameters * params, bool listKernel);
void common_free(Common * c);
void common_graph_drop(Common * c);
// Launch shape of a streaming step kernel over `threads` threads.  With the
// cap, blocks of 512 threads request just over half the CU's LDS, so one
// block -- 8 wavefronts -- runs per CU: fewer concurrent streams, measured
// 6.0 TB/s for the even step against 5.4 TB/s at full occupancy.  The LDS is
// reserved, never used.  Lattices too small for 16 such blocks per CU keep
// the plain launch (measured faster at 64^3).
struct StreamShape {
	int block;
	size_t lds;
};
StreamShape stream_shape(const Common * c, int defaultBlock, uint64_t threads);
// Grows the per-launch event pool to hold `launches` launches.
bool common_events(Common * c, int launches);
// Fills c->t from the recorded events (stream must be synchronised).
void common_collect(Common * c, int launches, bool perLaunch);

// Runs the step loop of `steps` steps for (omega, xForce) on c->stream by
// replaying a HIP graph; `record` issues the loop's kernel launches and is
// called, under stream capture, only when no graph was captured for these
// values yet.  The graph holds kernel launches only.
template <class F>
bool common_graph_run(Common * c, int steps, double omega, double xForce, F record)
{
	if (!c->graphExec || c->graphSteps != steps || c->graphOmega != omega || c->graphForce != xForce) {
		common_graph_drop(c);
		(void)hipGetLastError();
		if (!hip_ok(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return false;
		record();
		const hipError_t launchErr = hipGetLastError();
		const hipError_t endErr = hipStreamEndCapture(c->stream, &c->graph);
		if (!hip_ok(launchErr, "step kernel launch") || !hip_ok(endErr, "hipStreamEndCapture")) {
			common_graph_drop(c);
			return false;
		}
		if (!hip_ok(hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
			common_graph_drop(c);
			return false;
		}
		c->graphSteps = steps;
		c->graphOmega = omega;
		c->graphForce = xForce;
	}
	return hip_ok(hipGraphLaunch(c->graphExec, c->stream), "hipGraphLaunch");
}

// Pinned host buffer (hipHostMalloc) -- used for the host shadow.
void * host_alloc(size_t bytes);
void host_free(void * p);

}  // namespace lbm
