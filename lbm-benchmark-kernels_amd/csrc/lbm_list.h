// lbm_list.h -- list-aa-pv-soa state shared by the single-lattice kernel
// (lbm_list.hip) and the multi-lattice batch (lbm_batch.hip).
#pragma once

#include "lbm_device.h"
#include "lbm_runtime.h"

namespace lbm {

struct ListData {
	KernelData kd;              // first: callers cast KernelData* to the kernel's struct
	Common c;
	int dims[3] = {0, 0, 0};
	int periodic[3] = {0, 0, 0};
	int blk[3] = {0, 0, 0};
	int nFluid = 0;
	int nCells = 0;
	int adjStride = 0;
	int nRuns = 0;                // runs of consecutive adjacency (no forced breaks)
	int nRunsWave = 0;            // runs with a forced break at every 64-node wavefront
	int nWaves = 0;
	int runStride = 0;
	unsigned long long * dRunMask = nullptr;   // per wavefront: bit l = node 64w+l starts a run
	uint32_t * dRunBase = nullptr;             // per wavefront: index of its first run
	uint32_t * dRunAdj = nullptr;              // [18][runStride]: adjacency of each run's first node
	uint32_t * dGrid = nullptr;
	uint32_t * dCoords = nullptr;
	uint32_t * dAdj = nullptr;
	void * dPdf = nullptr;       // 19 * nCells values of the storage type
	void * hPdf = nullptr;       // host shadow, same layout and type, == kd.Pdfs[0]
	double * dMoments = nullptr; // reduction scratch: 2 per block + 2
	uint32_t * dOddOrder = nullptr; // odd step, -odd-order band: node block of each launched block
	uint32_t oddGrid = 0;           // blocks launched for the odd step
	int momentBlocks = 0;
	uint32_t * hGrid = nullptr;  // cell -> node id, for GetNode/SetNode
};

// The kernel's state of a KernelData made by D3Q19ListAaPvInit_PushSoAHip,
// or nullptr.
ListData * list_data(KernelData * kd);

// Host shadow -> device when SetNode touched it; device -> host shadow.
bool list_upload(ListData * L);
bool list_download(ListData * L);
// Bytes of one stored value for a -prec setting.
size_t elem_bytes(int prec);

// Host copy element <-> node-view value (f) for a -prec setting: FP64
// values as stored, shifted storage as w_d + stored (FP16: scaled by 2^-k).
double shadow_value(int prec, int f16Shift, const void * h, size_t i, int d);
void shadow_store(int prec, int f16Shift, void * h, size_t i, int d, double v);
// Nodes per thread of the even step (1, or 2 / 4 for FP32 / FP16 storage).
int list_even_vec(const ListData * L);
// Lattice weights w_d in direction order.
extern const double kW[19];

// Adds nBlocks (mass, x momentum) pairs in a fixed order (one block of B
// threads): the second pass of the FP64 moment sums.
template <int B>
__global__ __launch_bounds__(B) void k_sum_pairs(const double * __restrict__ partial, uint32_t nBlocks,
                                                 double * __restrict__ out)
{
	__shared__ double sm[B], sj[B];
	double m = 0.0, jx = 0.0;
	for (uint32_t i = threadIdx.x; i < nBlocks; i += B) {
		m += partial[2 * i];
		jx += partial[2 * i + 1];
	}
	sm[threadIdx.x] = m;
	sj[threadIdx.x] = jx;
	__syncthreads();
	for (int off = B / 2; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) {
			sm[threadIdx.x] += sm[threadIdx.x + off];
			sj[threadIdx.x] += sj[threadIdx.x + off];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		out[0] = sm[0];
		out[1] = sj[0];
	}
}

// aa-soa counterparts of the list kernel's moment sums and step runs
// (lbm_aa.hip): false with the error recorded.
bool aa_moments(KernelData * kd, double * mass, double * jx);
bool aa_steps(KernelData * kd, const CaseData * cd);
bool aa_download_shadow(KernelData * kd);

// Adjacency of node n for the odd step, from the full table or the runs.
struct AdjFull {
	const uint32_t * adj;
	uint32_t stride;
	__device__ __forceinline__ void fetch(uint32_t n, uint32_t (&a)[18]) const
	{
#pragma unroll
		for (int d = 0; d < 18; ++d) a[d] = adj[(size_t)d * stride + n];
	}
};

struct AdjRuns {
	const unsigned long long * mask;
	const uint32_t * base;
	const uint32_t * runAdj;
	uint32_t runStride;
	__device__ __forceinline__ void fetch(uint32_t n, uint32_t (&a)[18]) const
	{
		const uint32_t w = n >> 6;
		const int lane = (int)(n & 63u);
		// run starts at or before this node; lane 0 of a wavefront always starts one
		const unsigned long long upTo = mask[w] & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
		const uint32_t r = base[w] + (uint32_t)__popcll(upTo) - 1u;
		const uint32_t off = (uint32_t)(lane - (63 - __clzll(upTo)));
#pragma unroll
		for (int d = 0; d < 18; ++d) a[d] = runAdj[(size_t)d * runStride + r] + off;
	}
};

// Block order of the odd step: the device hands consecutive blocks to its
// 8 XCDs in turn; here every XCD takes kXcdGroup consecutive node blocks
// before the next XCD does, so the partly shared cache lines of neighbouring
// blocks (runs of ~5 nodes cut the neighbour planes into pieces) meet in one
// L2 more often: 2.4 % less odd-step time at 256^3 porous, measured over
// 1..64.  The grid is padded to a multiple of 8 * kXcdGroup blocks.
constexpr uint32_t kXcdGroup = 8;

__device__ __forceinline__ uint32_t xcd_grouped_block(uint32_t b)
{
	const uint32_t x = b % 8, j = b / 8;
	return ((j / kXcdGroup) * 8 + x) * kXcdGroup + j % kXcdGroup;
}

// The same order with 2^gShift consecutive blocks per XCD turn (the grid a
// multiple of 8 << gShift).
__device__ __forceinline__ uint32_t xcd_grouped_block(uint32_t b, uint32_t gShift)
{
	const uint32_t x = b % 8, j = b / 8;
	return (((j >> gShift) * 8 + x) << gShift) + (j & ((1u << gShift) - 1u));
}

// ------------------------------------------------------ per-node steps
// Even step: node n loads its 19 local values, collides and stores
// direction d into its own slot inv(d) (src/BenchKernelD3Q19ListAaPv.c:380-465).
// FORCE_LOAD adds the body-force increment of the preceding odd step to each
// loaded value.  NT streams the lattice with non-temporal loads and stores:
// every value is touched once per step, so caching it only evicts useful
// lines.  Codec: storage type and compute precision (FP64 values, or
// deviations f - w in FP32 / scaled FP16 computed in FP32).
template <class Codec, bool FORCE_LOAD, bool NT>
__device__ __forceinline__ void even_node(typename Codec::S * pdf, Codec cod, uint32_t n, uint32_t nCells,
                                          const TrtC<typename Codec::R> & k, const ForceT<typename Codec::R> & ft)
{
	using R = typename Codec::R;
	R f[19];
#pragma unroll
	for (int d = 0; d < 19; ++d) {
		f[d] = cod.template load<NT>(&pdf[(size_t)d * nCells + n]);
		if (FORCE_LOAD) f[d] = f[d] + ft.f[d];
	}
	trt_collide(f, k);
#pragma unroll
	for (int d = 0; d < 19; ++d) cod.template store<NT>(&pdf[(size_t)inv_dir(d) * nCells + n], f[d]);
}

// Even step over V consecutive nodes n0 .. n0+V-1, for the narrow storage
// types: one 8-byte access per direction carries V values (V = 2 for FP32,
// 4 for FP16), so a wavefront moves as many bytes per instruction as in
// FP64.  Needs nCells % V == 0 and n0 % V == 0 (V-aligned accesses); slots
// past nFluid in the last group are read (padding, zero-filled at Init) but
// not written.  Each node's arithmetic is that of even_node.
template <class Codec, int V, bool FORCE_LOAD, bool NT>
__device__ __forceinline__ void even_nodes(typename Codec::S * pdf, Codec cod, uint32_t n0, uint32_t nFluid,
                                           uint32_t nCells, const TrtC<typename Codec::R> & k,
                                           const ForceT<typename Codec::R> & ft)
{
	using R = typename Codec::R;
	using S = typename Codec::S;
	using VT = VecS<S, V>;
	R f[V][19];
#pragma unroll
	for (int d = 0; d < 19; ++d) {
		const VT x = ld<NT>(reinterpret_cast<const VT *>(&pdf[(size_t)d * nCells + n0]));
#pragma unroll
		for (int j = 0; j < V; ++j) {
			f[j][d] = cod.dec(x[j]);
			if (FORCE_LOAD) f[j][d] = f[j][d] + ft.f[d];
		}
	}
#pragma unroll
	for (int j = 0; j < V; ++j) trt_collide(f[j], k);
	if (n0 + V <= nFluid) {
#pragma unroll
		for (int d = 0; d < 19; ++d) {
			VT y;
#pragma unroll
			for (int j = 0; j < V; ++j) y[j] = cod.enc(f[j][d]);
			st<NT>(reinterpret_cast<VT *>(&pdf[(size_t)inv_dir(d) * nCells + n0]), y);
		}
	} else {
		const uint32_t m = nFluid - n0;
#pragma unroll
		for (int d = 0; d < 19; ++d)
#pragma unroll
			for (int j = 0; j < V; ++j)
				if ((uint32_t)j < m) cod.template store<NT>(&pdf[(size_t)inv_dir(d) * nCells + n0 + j], f[j][d]);
	}
}

// Odd step of node n with adjacency a (src/BenchKernelD3Q19ListAaPv.c:755-845):
// direction d arrives from slot a[inv d] (the neighbour's swapped store of
// the even step, or the node's own slot d for a bounce-back link); the
// post-collision value goes to a[d].  Every slot is read and written by
// exactly one node, so the update is in place and race free.  FORCE_STORE
// adds the force increment of this step to each stored value: a value
// stored through a bounce-back entry is read back by the same node as
// direction inv(d), any other by the neighbour as d.
// Node block of launched block b in the odd step: in launch order (plain),
// XCD-grouped, or from a table (band: every XCD takes the blocks of one
// eighth of the y range, so neighbours in x and y mostly meet in its L2).
struct OddOrder {
	const uint32_t * table;    // band order, or nullptr
	uint32_t grouped;          // 1: xcd_grouped_block
	__device__ __forceinline__ uint32_t block(uint32_t b) const
	{
		return table ? table[b] : grouped ? xcd_grouped_block(b) : b;
	}
};

template <class Codec, bool FORCE_LOAD, bool FORCE_STORE>
__device__ __forceinline__ void odd_node(typename Codec::S * pdf, Codec cod, const uint32_t (&a)[18], uint32_t n,
                                         uint32_t nCells, const TrtC<typename Codec::R> & k,
                                         const ForceT<typename Codec::R> & ft)
{
	using R = typename Codec::R;
	R f[19];
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		f[d] = cod.template load<false>(&pdf[a[inv_dir(d)]]);
		if (FORCE_LOAD) f[d] = f[d] + ft.f[d];
	}
	const size_t cSlot = (size_t)dC * nCells + n;
	f[dC] = cod.template load<false>(&pdf[cSlot]);
	if (FORCE_LOAD) f[dC] = f[dC] + ft.f[dC];
	trt_collide(f, k);
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		R v = f[d];
		if (FORCE_STORE) {
			const bool bounce = a[d] == (uint32_t)inv_dir(d) * nCells + n;
			v = v + (bounce ? ft.f[inv_dir(d)] : ft.f[d]);
		}
		cod.template store<false>(&pdf[a[d]], v);
	}
	cod.template store<false>(&pdf[cSlot], FORCE_STORE ? f[dC] + ft.f[dC] : f[dC]);
}

}  // namespace lbm
