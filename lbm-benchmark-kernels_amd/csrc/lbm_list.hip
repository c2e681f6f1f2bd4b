// lbm_list.hip -- list-aa-pv-soa on MI355X: fluid-node list, push adjacency
// and the even/odd AA steps as gfx950 kernels, behind the reference's
// kernel-plugin ABI.
//
// Reference behaviour reproduced here:
//   * node numbering and adjacency values of
//     src/BenchKernelD3Q19ListAaPvCommon.c:627-793, bit for bit;
//   * padded SoA PDF layout d * nCells + n (src/BenchKernelD3Q19ListAaCommon.h:79-103)
//     with nCells from the reference's padding rule (src/Padding.c:195-244);
//   * even step: local load, TRT collision, store of direction d into slot
//     inv(d) of the same node (src/BenchKernelD3Q19ListAaPv.c:380-465);
//   * odd step: direction d loaded from adj[n][inv d], stored to adj[n][d]
//     (src/BenchKernelD3Q19ListAaPv.c:755-845);
//   * VERIFICATION builds add the Luo force in node view after every step
//     (src/BenchKernelD3Q19ListAaPv.c:100-126, src/Kernel.c:759-803).
//     Here the increment of each step is added where the next step loads
//     the value, and after the last step where it is stored -- the same
//     single addition on the same operands, without an extra lattice pass.
//
// Device layout: PDFs as the reference (19 planes of nCells doubles; nCells
// * 8 B is a multiple of 128 B after padding, so every plane starts on a
// 128-B line; with -prec fp32s / fp16s the planes hold FP32 / FP16
// deviations f - w instead).  Adjacency transposed to [18][adjStride] uint32
// so that lane n of a wavefront reads entry d of node n next to lane n+1;
// the odd step reads a run-compressed copy of it (per 64-node wavefront a
// run-start mask and base, per run the 18 entries of its first node) and
// visits its blocks in an XCD-aware order (OddOrder).
#include "lbm_list.h"

// The program's output functions (the reference's Vtk.o / Kernel.o, or
// liblbmhost.so); weak, so the library loads without them.
extern "C" {
void VtkWrite(LatticeDesc * ld, KernelData * kd, CaseData * cd, int iteration) __attribute__((weak));
void KernelStatistics(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration) __attribute__((weak));
}

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <new>
#include <vector>

namespace lbm {

static ListData * as_list(KernelData * kd) { return reinterpret_cast<ListData *>(kd); }
static const ListData * as_list(const KernelData * kd) { return reinterpret_cast<const ListData *>(kd); }

// ------------------------------------------------------------ list build
// Node numbering visits blocks (bx, by, bz) and inside each block x, y, z
// with z fastest (src/BenchKernelD3Q19ListAaPvCommon.c:631-656).  trav_xyz
// maps a position t of that traversal to its cell; the node id of a fluid
// cell is the number of fluid cells before it in the traversal, computed
// with wavefront ballots and a prefix sum over block totals.
struct NumGeom {
	int X, Y, Z, bx, by, bz;
};

__device__ __forceinline__ void trav_xyz(uint32_t t, const NumGeom & g, int & x, int & y, int & z)
{
	const uint32_t slab = (uint32_t)g.bx * g.Y * g.Z;            // every x block before the last is full
	const int i = (int)(t / slab);
	uint32_t r = t - (uint32_t)i * slab;
	const int ex = min(g.bx, g.X - i * g.bx);
	const uint32_t rowBlk = (uint32_t)g.by * ex * g.Z;
	const int j = (int)(r / rowBlk);
	r -= (uint32_t)j * rowBlk;
	const int ey = min(g.by, g.Y - j * g.by);
	const uint32_t cellBlk = (uint32_t)ex * ey * g.bz;
	const int k = (int)(r / cellBlk);
	r -= (uint32_t)k * cellBlk;
	const int ez = min(g.bz, g.Z - k * g.bz);
	const uint32_t plane = (uint32_t)ey * ez;
	const int lx = (int)(r / plane);
	r -= (uint32_t)lx * plane;
	const int ly = (int)(r / ez);
	const int lz = (int)(r - (uint32_t)ly * ez);
	x = i * g.bx + lx;
	y = j * g.by + ly;
	z = k * g.bz + lz;
}

constexpr int kNumBlock = 256;   // 4 wavefronts

__global__ __launch_bounds__(kNumBlock) void k_num_count(const int * __restrict__ mask, NumGeom g,
                                                         uint32_t N, uint32_t * __restrict__ blockSums)
{
	const uint32_t t = blockIdx.x * kNumBlock + threadIdx.x;
	int fluid = 0;
	if (t < N) {
		int x, y, z;
		trav_xyz(t, g, x, y, z);
		fluid = mask[((size_t)z * g.Y + y) * g.X + x] != LBM_CELL_OBSTACLE;
	}
	const unsigned long long b = __ballot(fluid);
	__shared__ uint32_t waveCount[kNumBlock / 64];
	if (__lane_id() == 0) waveCount[threadIdx.x / 64] = (uint32_t)__popcll(b);
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t s = 0;
		for (int w = 0; w < kNumBlock / 64; ++w) s += waveCount[w];
		blockSums[blockIdx.x] = s;
	}
}

// Exclusive scan of n block totals in one workgroup; total -> *sum.
__global__ __launch_bounds__(1024) void k_scan_once(uint32_t * __restrict__ v, uint32_t n, uint32_t * __restrict__ sum)
{
	__shared__ uint32_t part[1024];
	const uint32_t tid = threadIdx.x;
	const uint32_t per = (n + 1023) / 1024;
	const uint32_t s = min(n, tid * per), e = min(n, s + per);
	uint32_t acc = 0;
	for (uint32_t i = s; i < e; ++i) acc += v[i];
	part[tid] = acc;
	__syncthreads();
	for (uint32_t off = 1; off < 1024; off <<= 1) {
		const uint32_t add = tid >= off ? part[tid - off] : 0;
		__syncthreads();
		part[tid] += add;
		__syncthreads();
	}
	uint32_t run = tid ? part[tid - 1] : 0;
	for (uint32_t i = s; i < e; ++i) {
		const uint32_t c = v[i];
		v[i] = run;
		run += c;
	}
	if (tid == 1023) *sum = part[1023];
}

__global__ __launch_bounds__(kNumBlock) void k_num_assign(const int * __restrict__ mask, NumGeom g, uint32_t N,
                                                          const uint32_t * __restrict__ blockOffset,
                                                          uint32_t * __restrict__ grid, uint32_t * __restrict__ coords)
{
	const uint32_t t = blockIdx.x * kNumBlock + threadIdx.x;
	int x = 0, y = 0, z = 0, fluid = 0;
	size_t cell = 0;
	if (t < N) {
		trav_xyz(t, g, x, y, z);
		cell = ((size_t)z * g.Y + y) * g.X + x;
		fluid = mask[cell] != LBM_CELL_OBSTACLE;
	}
	const unsigned long long b = __ballot(fluid);
	const int lane = __lane_id();
	const uint32_t below = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
	__shared__ uint32_t waveCount[kNumBlock / 64];
	if (lane == 0) waveCount[threadIdx.x / 64] = (uint32_t)__popcll(b);
	__syncthreads();
	uint32_t waveBase = 0;
	for (int w = 0; w < (int)(threadIdx.x / 64); ++w) waveBase += waveCount[w];
	if (t < N) {
		if (fluid) {
			const uint32_t id = blockOffset[blockIdx.x] + waveBase + below;
			grid[cell] = id;
			coords[3 * (size_t)id + 0] = (uint32_t)x;
			coords[3 * (size_t)id + 1] = (uint32_t)y;
			coords[3 * (size_t)id + 2] = (uint32_t)z;
		} else {
			grid[cell] = 0xffffffffu;
		}
	}
}

struct AdjGeom {
	int X, Y, Z, px, py, pz;
	uint32_t nFluid, nCells, stride;
};

// Push adjacency of src/BenchKernelD3Q19ListAaPvCommon.c:700-793: the slot
// d * nCells + m of the neighbour m in direction d, or the node's own slot
// inv(d) * nCells + n (half-way bounce-back) when that neighbour is solid or
// outside a non-periodic face.  Once a coordinate leaves the box through a
// periodic face, every out-of-range coordinate is wrapped (:729-761).
__global__ __launch_bounds__(256) void k_adjacency(const int * __restrict__ mask, const uint32_t * __restrict__ grid,
                                                   const uint32_t * __restrict__ coords, AdjGeom g,
                                                   uint32_t * __restrict__ adj)
{
	const uint32_t n = blockIdx.x * 256 + threadIdx.x;
	if (n >= g.nFluid) return;
	const int x = (int)coords[3 * (size_t)n + 0];
	const int y = (int)coords[3 * (size_t)n + 1];
	const int z = (int)coords[3 * (size_t)n + 2];
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		int nx = x + cx(d), ny = y + cy(d), nz = z + cz(d);
		const bool ox = nx < 0 || nx >= g.X, oy = ny < 0 || ny >= g.Y, oz = nz < 0 || nz >= g.Z;
		const uint32_t bounce = (uint32_t)inv_dir(d) * g.nCells + n;
		uint32_t v;
		if ((ox && g.px) || (oy && g.py) || (oz && g.pz)) {
			nx = nx < 0 ? g.X - 1 : (nx >= g.X ? 0 : nx);
			ny = ny < 0 ? g.Y - 1 : (ny >= g.Y ? 0 : ny);
			nz = nz < 0 ? g.Z - 1 : (nz >= g.Z ? 0 : nz);
			const size_t c = ((size_t)nz * g.Y + ny) * g.X + nx;
			v = mask[c] == LBM_CELL_OBSTACLE ? bounce : (uint32_t)d * g.nCells + grid[c];
		} else if (ox || oy || oz) {
			v = bounce;
		} else {
			const size_t c = ((size_t)nz * g.Y + ny) * g.X + nx;
			v = mask[c] == LBM_CELL_OBSTACLE ? bounce : (uint32_t)d * g.nCells + grid[c];
		}
		adj[(size_t)d * g.stride + n] = v;
	}
}

// Counts nodes n >= 1 whose 18 adjacency entries all equal those of n-1
// plus one, i.e. the starts of runs (src/BenchKernelD3Q19ListAaPvCommon.c:360-385).
__global__ __launch_bounds__(256) void k_run_breaks(const uint32_t * __restrict__ adj, uint32_t nFluid,
                                                    uint32_t stride, uint32_t * __restrict__ breaks)
{
	const uint32_t n = blockIdx.x * 256 + threadIdx.x;
	int br = 0;
	if (n >= 1 && n < nFluid) {
		for (int d = 0; d < 18; ++d)
			if (adj[(size_t)d * stride + n] != adj[(size_t)d * stride + n - 1] + 1u) { br = 1; break; }
	}
	const unsigned long long b = __ballot(br);
	if (__lane_id() == 0 && b) atomicAdd(breaks, (uint32_t)__popcll(b));
}

// Run-compressed adjacency for the odd step.  Inside a run of nodes whose
// 18 entries all grow by one from node to node -- the reference's
// "consecutive nodes" (src/BenchKernelD3Q19ListAaPvCommon.c:326-469) -- only
// the first node's entries are stored; the others add their offset in the
// run.  Runs are cut at every 64-node wavefront, so a wavefront finds its
// runs from one 64-bit start mask and one base index.
__global__ __launch_bounds__(256) void k_wave_runs(const uint32_t * __restrict__ adj, uint32_t nFluid, uint32_t stride,
                                                   uint32_t nWaves, unsigned long long * __restrict__ runMask,
                                                   uint32_t * __restrict__ runCount)
{
	const uint32_t n = blockIdx.x * 256 + threadIdx.x;
	int start = 0;
	if (n < nFluid) {
		start = (n & 63u) == 0;
		if (!start)
			for (int d = 0; d < 18; ++d)
				if (adj[(size_t)d * stride + n] != adj[(size_t)d * stride + n - 1] + 1u) { start = 1; break; }
	}
	const unsigned long long b = __ballot(start);
	const uint32_t w = n >> 6;
	if (__lane_id() == 0 && w < nWaves) {
		runMask[w] = b;
		runCount[w] = (uint32_t)__popcll(b);
	}
}

__global__ __launch_bounds__(256) void k_run_compact(const uint32_t * __restrict__ adj, uint32_t nFluid, uint32_t stride,
                                                     const unsigned long long * __restrict__ runMask,
                                                     const uint32_t * __restrict__ runBase, uint32_t runStride,
                                                     uint32_t * __restrict__ runAdj)
{
	const uint32_t n = blockIdx.x * 256 + threadIdx.x;
	if (n >= nFluid) return;
	const uint32_t w = n >> 6, lane = n & 63u;
	const unsigned long long m = runMask[w];
	if (!((m >> lane) & 1ull)) return;
	const uint32_t r = runBase[w] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
	for (int d = 0; d < 18; ++d) runAdj[(size_t)d * runStride + r] = adj[(size_t)d * stride + n];
}

// ---------------------------------------------------------------- steps
// The per-node bodies (even_node, even_nodes, odd_node) live in lbm_list.h,
// shared with the multi-lattice batch kernels.

// Node block of the even step: ascending (default), or (down) descending,
// so that the step starts on the highest nodes -- the x planes every XCD's
// y band of the preceding odd step ends on, whose lines may still be in the
// last-level cache -- and ends on the lowest, where the next odd step
// begins.  At 256^3 porous FP64 descending measured 4 % faster on one
// device and 0.5 % slower on another (eight alternating runs each).
__device__ __forceinline__ uint32_t even_block(uint32_t down)
{
	return down ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
}

template <class Codec, bool FORCE_LOAD, bool NT>
__global__ __launch_bounds__(512) void k_list_even(typename Codec::S * pdf, Codec cod, uint32_t nFluid, uint32_t nCells,
                                                   TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft, uint32_t down)
{
	const uint32_t n = even_block(down) * blockDim.x + threadIdx.x;
	if (n >= nFluid) return;
	even_node<Codec, FORCE_LOAD, NT>(pdf, cod, n, nCells, k, ft);
}

template <class Codec, int V, bool FORCE_LOAD, bool NT>
__global__ __launch_bounds__(512) void k_list_even_v(typename Codec::S * pdf, Codec cod, uint32_t nFluid, uint32_t nCells,
                                                     TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft,
                                                     uint32_t down)
{
	const uint32_t n0 = (even_block(down) * blockDim.x + threadIdx.x) * V;
	if (n0 >= nFluid) return;
	even_nodes<Codec, V, FORCE_LOAD, NT>(pdf, cod, n0, nFluid, nCells, k, ft);
}

template <class Codec, bool FORCE_LOAD, bool FORCE_STORE, class Adj>
__global__ __launch_bounds__(512) void k_list_odd(typename Codec::S * pdf, Codec cod, Adj adjacency, OddOrder order,
                                                  uint32_t nFluid, uint32_t nCells, TrtC<typename Codec::R> k,
                                                  ForceT<typename Codec::R> ft)
{
	const uint32_t blk = order.block(blockIdx.x);
	if (blk == 0xffffffffu) return;             // padding of the band table
	const uint32_t n = blk * blockDim.x + threadIdx.x;
	if (n >= nFluid) return;
	uint32_t a[18];
	adjacency.fetch(n, a);
	odd_node<Codec, FORCE_LOAD, FORCE_STORE>(pdf, cod, a, n, nCells, k, ft);
}

// Node sums in the local layout (after a completed Kernel() call), in FP64:
// per block the sum over its nodes of sum_d v_d and of sum_d c_x,d v_d, v
// the stored value (f, or the deviation f - w for shifted storage); a second
// launch adds the block sums in a fixed order, so results do not depend on
// timing.
template <class Codec>
__global__ __launch_bounds__(256) void k_list_moments(const typename Codec::S * pdf, Codec cod, uint32_t nFluid,
                                                      uint32_t nCells, double * __restrict__ partial)
{
	const uint32_t n = blockIdx.x * 256 + threadIdx.x;
	double m = 0.0, jx = 0.0;
	if (n < nFluid) {
		for (int d = 0; d < 19; ++d) {
			const double v = (double)cod.template load<false>(&pdf[(size_t)d * nCells + n]);
			m += v;
			jx += cx(d) * v;
		}
	}
	__shared__ double sm[256], sj[256];
	sm[threadIdx.x] = m;
	sj[threadIdx.x] = jx;
	__syncthreads();
	for (int off = 128; off > 0; off >>= 1) {
		if ((int)threadIdx.x < off) {
			sm[threadIdx.x] += sm[threadIdx.x + off];
			sj[threadIdx.x] += sj[threadIdx.x + off];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		partial[2 * blockIdx.x] = sm[0];
		partial[2 * blockIdx.x + 1] = sj[0];
	}
}

// --------------------------------------------------------- node access
// After any completed Kernel() call (always an even number of steps) and
// before the first one, the lattice is in the local layout: node n's value
// of direction d sits in slot d * nCells + n
// (src/BenchKernelD3Q19ListAaPvCommon.c:150-156).
static bool list_local_layout(const ListData * L)
{
	if (L->c.iteration % 2 != 0) return true;
	return fail("list-aa-pv-soa: node access after an even step is not supported by this build.");
}

// Shadow element <-> node-view value.  Shifted storage holds f - w_d
// (w_d = w_inv(d), so a slot keeps the same shift whichever node reads it).
const double kW[19] = {
	1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
	1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
	1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0 };

size_t elem_bytes(int prec) { return prec == PREC_FP64 ? 8 : prec == PREC_FP32S ? 4 : 2; }

double shadow_value(int prec, int f16Shift, const void * h, size_t i, int d)
{
	switch (prec) {
		case PREC_FP32S: return kW[d] + (double)((const float *)h)[i];
		case PREC_FP16S: return kW[d] + (double)(f16_to_f32(((const uint16_t *)h)[i]) / (float)(1 << f16Shift));
		default: return ((const double *)h)[i];
	}
}

void shadow_store(int prec, int f16Shift, void * h, size_t i, int d, double v)
{
	switch (prec) {
		case PREC_FP32S: ((float *)h)[i] = (float)(v - kW[d]); break;
		case PREC_FP16S: ((uint16_t *)h)[i] = f32_to_f16((float)(v - kW[d]) * (float)(1 << f16Shift)); break;
		default: ((double *)h)[i] = v;
	}
}

static double shadow_get(const ListData * L, size_t i, int d)
{
	return shadow_value(L->c.args.prec, L->c.args.f16Shift, L->hPdf, i, d);
}

static void shadow_set(ListData * L, size_t i, int d, double v)
{
	shadow_store(L->c.args.prec, L->c.args.f16Shift, L->hPdf, i, d, v);
}

// Node id of a fluid cell, or false (with the error recorded) for a cell
// outside the box or a solid cell -- the reference asserts both only in
// debug builds (src/BenchKernelD3Q19ListAaPvCommon.c:133-134).
static bool list_node(const ListData * L, int x, int y, int z, uint32_t * n)
{
	if (!list_local_layout(L)) return false;
	if (!L->hPdf) return fail("list-aa-pv-soa: no host copy of the lattice yet (-shadow lazy); call LbmHipSyncShadow first.");
	if (x < 0 || y < 0 || z < 0 || x >= L->dims[0] || y >= L->dims[1] || z >= L->dims[2])
		return fail("list-aa-pv-soa: node (%d, %d, %d) outside the lattice.", x, y, z);
	*n = L->hGrid[((size_t)z * L->dims[1] + y) * L->dims[0] + x];
	if (*n == 0xffffffffu) return fail("list-aa-pv-soa: node (%d, %d, %d) is solid.", x, y, z);
	return true;
}

static void list_get_node(KernelData * kd, int x, int y, int z, PdfT * pdfs)
{
	ListData * L = as_list(kd);
	uint32_t n;
	if (!list_node(L, x, y, z, &n)) return;
	for (int d = 0; d < 19; ++d) pdfs[d] = shadow_get(L, (size_t)d * L->nCells + n, d);
}

static void list_set_node(KernelData * kd, int x, int y, int z, PdfT * pdfs)
{
	ListData * L = as_list(kd);
	uint32_t n;
	if (!list_node(L, x, y, z, &n) || !shadow_writable(&L->c)) return;
	for (int d = 0; d < 19; ++d) shadow_set(L, (size_t)d * L->nCells + n, d, pdfs[d]);
	L->c.shadowDirty = 1;
}

static void list_bc_get(KernelData * kd, int x, int y, int z, int dir, PdfT * pdf)
{
	ListData * L = as_list(kd);
	uint32_t n;
	if (!list_node(L, x, y, z, &n) || dir < 0 || dir >= 19) return;
	*pdf = shadow_get(L, (size_t)dir * L->nCells + n, dir);
}

static void list_bc_set(KernelData * kd, int x, int y, int z, int dir, PdfT pdf)
{
	ListData * L = as_list(kd);
	uint32_t n;
	if (!list_node(L, x, y, z, &n) || dir < 0 || dir >= 19 || !shadow_writable(&L->c)) return;
	shadow_set(L, (size_t)dir * L->nCells + n, dir, pdf);
	L->c.shadowDirty = 1;
}

// ------------------------------------------------------------- Kernel()
bool list_upload(ListData * L)
{
	if (!L->c.shadowDirty) { L->c.t.uploadMs = 0.0; return true; }
	const double t0 = (double)clock() / CLOCKS_PER_SEC;
	if (!hip_ok(hipMemcpyAsync(L->dPdf, L->hPdf, elem_bytes(L->c.args.prec) * 19 * (size_t)L->nCells,
	                           hipMemcpyHostToDevice, L->c.stream), "upload of the host lattice")) return false;
	if (!hip_ok(hipStreamSynchronize(L->c.stream), "hipStreamSynchronize")) return false;
	L->c.t.uploadMs = ((double)clock() / CLOCKS_PER_SEC - t0) * 1e3;
	L->c.shadowDirty = 0;
	return true;
}

// The host copy, allocated here on first use with -shadow lazy.
static bool ensure_shadow(ListData * L)
{
	if (L->hPdf) return true;
	const size_t bytes = elem_bytes(L->c.args.prec) * 19 * (size_t)L->nCells;
	L->hPdf = host_alloc(bytes);
	if (!L->hPdf) return fail("allocating host lattice failed: %zu bytes.", bytes);
	L->kd.Pdfs[0] = (PdfT *)L->hPdf;
	L->kd.PdfsActive = (PdfT *)L->hPdf;
	return true;
}

bool list_download(ListData * L)
{
	if (L->c.shadowDirty) return fail("host lattice modified through SetNode since the last Kernel() call; not overwriting it.");
	if (!ensure_shadow(L)) return false;
	if (!hip_ok(hipMemcpyAsync(L->hPdf, L->dPdf, elem_bytes(L->c.args.prec) * 19 * (size_t)L->nCells,
	                           hipMemcpyDeviceToHost, L->c.stream), "download of the device lattice")) return false;
	if (!hip_ok(hipStreamSynchronize(L->c.stream), "hipStreamSynchronize")) return false;
	L->c.shadowStale = 0;
	return true;
}

// Algorithmic bytes per fluid node of the odd step: 19 loads and 19 stores
// of 8 B plus the adjacency -- 18 x 4 B per node in full, or per run 18 x
// 4 B plus 12 B of mask and base per 64-node wavefront.
static double list_odd_bytes(const ListData * L)
{
	const double pdf = 2.0 * 19 * (double)elem_bytes(L->c.args.prec);
	if (!L->c.args.adjRle || L->nFluid == 0) return pdf + 18 * 4.0;
	return pdf + (72.0 * L->nRunsWave + 12.0 * L->nWaves) / L->nFluid;
}

static double list_even_bytes(const ListData * L) { return 2.0 * 19 * (double)elem_bytes(L->c.args.prec); }

// Nodes per thread of the even step: 8-byte accesses for the narrow storage
// types when the plane stride allows aligned vectors (-vec 1 turns it off).
int list_even_vec(const ListData * L)
{
	const int v = L->c.args.prec == PREC_FP32S ? 2 : L->c.args.prec == PREC_FP16S ? 4 : 1;
	if (!L->c.args.evenVec || L->nCells % v != 0) return 1;
	return v;
}

struct LaunchCfg {
	dim3 eblock;
	size_t elds;
	dim3 grid, block;
	OddOrder order;
	hipStream_t stream;
	uint32_t nF, nC;
	bool rle, nt;
	uint32_t down;          // even step in descending block order
	// blocks of the even step when each thread takes v nodes
	unsigned evenThreads(int v) const
	{
		const uint32_t threads = (nF + (uint32_t)v - 1) / (uint32_t)v;
		return (threads + eblock.x - 1) / eblock.x;
	}
};

template <class Codec, int V>
static void launch_even(const LaunchCfg & c, typename Codec::S * pdf, Codec cod, const TrtConsts & k0,
                        const ForceTerms & f0, bool fl)
{
	using R = typename Codec::R;
	const TrtC<R> k = trt_cast<R>(k0);
	const ForceT<R> ft = force_cast<R>(f0);
	if constexpr (V == 1) {
		const dim3 g(c.evenThreads(1));
		if (fl && c.nt)  k_list_even<Codec, true, true><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else if (fl)     k_list_even<Codec, true, false><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else if (c.nt)   k_list_even<Codec, false, true><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else             k_list_even<Codec, false, false><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
	} else {
		const dim3 g(c.evenThreads(V));
		if (fl && c.nt)  k_list_even_v<Codec, V, true, true><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else if (fl)     k_list_even_v<Codec, V, true, false><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else if (c.nt)   k_list_even_v<Codec, V, false, true><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
		else             k_list_even_v<Codec, V, false, false><<<g, c.eblock, c.elds, c.stream>>>(pdf, cod, c.nF, c.nC, k, ft, c.down);
	}
}

template <class Codec, class Adj>
static void launch_odd_adj(const LaunchCfg & c, typename Codec::S * pdf, Codec cod, Adj adj, const TrtC<typename Codec::R> & k,
                           const ForceT<typename Codec::R> & ft, bool fl, bool fs)
{
	if (fs)      k_list_odd<Codec, true, true><<<c.grid, c.block, 0, c.stream>>>(pdf, cod, adj, c.order, c.nF, c.nC, k, ft);
	else if (fl) k_list_odd<Codec, true, false><<<c.grid, c.block, 0, c.stream>>>(pdf, cod, adj, c.order, c.nF, c.nC, k, ft);
	else         k_list_odd<Codec, false, false><<<c.grid, c.block, 0, c.stream>>>(pdf, cod, adj, c.order, c.nF, c.nC, k, ft);
}

template <class Codec>
static void launch_odd(const LaunchCfg & c, typename Codec::S * pdf, Codec cod, const AdjFull & full, const AdjRuns & runs,
                       const TrtConsts & k0, const ForceTerms & f0, bool fl, bool fs)
{
	using R = typename Codec::R;
	const TrtC<R> k = trt_cast<R>(k0);
	const ForceT<R> ft = force_cast<R>(f0);
	if (c.rle) launch_odd_adj(c, pdf, cod, runs, k, ft, fl, fs);
	else       launch_odd_adj(c, pdf, cod, full, k, ft, fl, fs);
}

// Output during the run, as the reference's VTK_OUTPUT / STATISTICS builds
// write it from inside the kernel (src/BenchKernelD3Q19ListAaPv.c:67-77,
// 128-141, 146-156): before the first step VtkWrite(-1) and
// KernelStatistics(0); after the step pair starting at iteration it,
// VtkWrite(it) when it % VtkModulus == 0 and KernelStatistics(it); at the end
// VtkWrite(max) and KernelStatistics(max).  KernelStatistics writes only
// when it % 10 == 0 or it % StatisticsModulus == 0 (src/Kernel.c:633-646), so
// the lattice is copied to the host for those pairs only.  The functions are
// the program's own (the reference's Vtk.o / Kernel.o, or liblbmhost.so).
struct RunOutput {
	LatticeDesc * ld = nullptr;
	bool vtk = false, stats = false;
	bool at(const CaseData * cd, int it) const
	{
		return (vtk && it % cd->VtkModulus == 0) ||
		       (stats && (it % 10 == 0 || it % cd->StatisticsModulus == 0));
	}
};

static bool run_output_setup(const ListData * L, LatticeDesc * ld, const CaseData * cd, RunOutput * o)
{
	o->ld = ld;
	if (!ld) return true;
	o->vtk = cd->VtkOutput != 0;
	o->stats = L->c.args.statistics != 0;
	if (o->vtk && !VtkWrite) return fail("VTK output requested but the program does not link VtkWrite.");
	if (o->stats && !KernelStatistics) return fail("-statistics requested but the program does not link KernelStatistics.");
	if (o->vtk && cd->VtkModulus <= 0) return fail("VtkModulus must be > 0.");
	if (o->stats && cd->StatisticsModulus <= 0) return fail("StatisticsModulus must be > 0.");
	return true;
}

// Host copy of the current state, for the output functions' GetNode.
static bool output_sync(ListData * L)
{
	if (L->c.shadowDirty || (L->hPdf && !L->c.shadowStale)) return true;   // the host copy is current
	return list_download(L);
}

static bool list_steps(KernelData * kd, const CaseData * cd, LatticeDesc * ld)
{
	ListData * L = as_list(kd);
	Common & c = L->c;
	const int steps = cd->MaxIterations > 0 ? cd->MaxIterations + (cd->MaxIterations & 1) : 0;
	if (!(cd->Omega > 0.0 && cd->Omega < 2.0)) { return fail("omega must be in (0, 2), got %f.", cd->Omega); }
	RunOutput out;
	if (!run_output_setup(L, ld, cd, &out)) return false;
	CaseData * ccd = const_cast<CaseData *>(cd);
	if (out.vtk || out.stats) {
		if (!output_sync(L)) return false;
		if (out.vtk) VtkWrite(ld, kd, ccd, -1);
		if (out.stats) KernelStatistics(kd, ld, ccd, 0);
	}
	if (!list_upload(L)) return false;

	const TrtConsts k = trt_consts(cd->Omega, c.args.magic);
	const ForceTerms ft = force_terms(cd->XForce);
	const bool force = c.args.forceOn != 0;
	const bool perLaunch = c.launchEvents != 0;
	if (perLaunch && !common_events(&c, steps)) return false;

	const int bs = c.args.blockSize;
	const dim3 grid(L->oddGrid), block((unsigned)bs);
	const OddOrder order = {L->dOddOrder, c.args.oddOrder == 1 ? 1u : 0u};
	const uint32_t nF = (uint32_t)L->nFluid, nC = (uint32_t)L->nCells;
	const AdjFull full = {L->dAdj, (uint32_t)L->adjStride};
	const AdjRuns runs = {L->dRunMask, L->dRunBase, L->dRunAdj, (uint32_t)L->runStride};
	const int vec = list_even_vec(L);
	// non-temporal even-step accesses by default only when what a step
	// touches -- the fluid nodes' PDFs and the odd step's adjacency --
	// exceeds the last-level cache (256^3 porous FP64: 0.69 against
	// 0.73 ms); a cache-sized lattice streams faster with plain accesses
	// (64^3 channel FP64 15.6 -> 18.4 GFLUP/s, 128^3 porous FP32 31.7 ->
	// 38.8; 128^3 porous FP64, 290 MB, 20.2 with and 17.4 without)
	const size_t touched = elem_bytes(c.args.prec) * 19 * (size_t)L->nFluid +
	                       (c.args.adjRle ? 72 * (size_t)L->nRunsWave + 12 * (size_t)L->nWaves : 72 * (size_t)L->nFluid);
	const bool nt = c.args.nontemporal == 1 || (c.args.nontemporal == -1 && touched > kLastLevelCacheBytes);
	// four FP16 nodes per thread run faster at full occupancy (measured +6 % at
	// 256^3 porous); the other storage types with the occupancy cap
	const StreamShape es = (vec == 4 && !c.args.evenCapGiven) ? StreamShape{bs, 0}
	                       : stream_shape(&c, bs, ((uint64_t)L->nFluid + vec - 1) / vec);
	const LaunchCfg cfg = {dim3((unsigned)es.block), es.lds, grid, block, order, c.stream, nF, nC,
	                       c.args.adjRle != 0, nt, c.args.evenDown ? 1u : 0u};
	const float f16s = (float)(1 << c.args.f16Shift);
	auto even = [&](bool fl) {
		switch (c.args.prec) {
			case PREC_FP32S:
				if (vec == 2) launch_even<CodecF32, 2>(cfg, (float *)L->dPdf, CodecF32{}, k, ft, fl);
				else          launch_even<CodecF32, 1>(cfg, (float *)L->dPdf, CodecF32{}, k, ft, fl);
				break;
			case PREC_FP16S:
				if (vec == 4) launch_even<CodecF16, 4>(cfg, (uint16_t *)L->dPdf, CodecF16{f16s, 1.0f / f16s}, k, ft, fl);
				else          launch_even<CodecF16, 1>(cfg, (uint16_t *)L->dPdf, CodecF16{f16s, 1.0f / f16s}, k, ft, fl);
				break;
			default: launch_even<CodecF64, 1>(cfg, (double *)L->dPdf, CodecF64{}, k, ft, fl);
		}
	};
	auto odd = [&](bool fl, bool fs) {
		switch (c.args.prec) {
			case PREC_FP32S: launch_odd(cfg, (float *)L->dPdf, CodecF32{}, full, runs, k, ft, fl, fs); break;
			case PREC_FP16S: launch_odd(cfg, (uint16_t *)L->dPdf, CodecF16{f16s, 1.0f / f16s}, full, runs, k, ft, fl, fs); break;
			default:         launch_odd(cfg, (double *)L->dPdf, CodecF64{}, full, runs, k, ft, fl, fs);
		}
	};
	(void)hipGetLastError();   // launch checks below must see only this loop's errors
	if (!hip_ok(hipEventRecord(c.evLoop[0], c.stream), "hipEventRecord")) return false;
	bool segmentStart = true;  // the force of the previous step is in the stored values
	bool synced = true;        // the host copy holds the state of the last launch
	const bool useGraph = c.args.graph && !perLaunch && !out.vtk && !out.stats && steps > 0 && L->nFluid > 0;
	if (useGraph) {
		// the same launches as the loop below without output, replayed
		auto record = [&]() {
			for (int it = 0; it < steps; it += 2) {
				even(force && it > 0);
				odd(force, force && it + 2 >= steps);
			}
		};
		if (!common_graph_run(&c, steps, cd->Omega, cd->XForce, record)) return false;
		synced = false;
	}
	for (int it = 0; it < steps && L->nFluid > 0 && !useGraph; it += 2) {
		const bool output = out.at(cd, it);
		synced = false;
		const bool last = it + 2 >= steps || output;
		if (perLaunch) (void)hipEventRecord(c.ev[2 * it], c.stream);
		even(force && !segmentStart);
		if (perLaunch) {
			(void)hipEventRecord(c.ev[2 * it + 1], c.stream);
			(void)hipEventRecord(c.ev[2 * it + 2], c.stream);
		}
		odd(force, force && last);
		if (perLaunch) (void)hipEventRecord(c.ev[2 * it + 3], c.stream);
		segmentStart = last;
		if (output) {
			if (!hip_ok(hipGetLastError(), "step kernel launch")) return false;
			c.iteration = it + 1;
			c.shadowStale = 1;
			if (!output_sync(L)) return false;
			synced = true;
			if (out.vtk && it % cd->VtkModulus == 0) VtkWrite(ld, kd, ccd, it);
			if (out.stats) KernelStatistics(kd, ld, ccd, it);
		}
	}
	if (!hip_ok(hipGetLastError(), "step kernel launch")) return false;
	if (!hip_ok(hipEventRecord(c.evLoop[1], c.stream), "hipEventRecord")) return false;
	if (!hip_ok(hipStreamSynchronize(c.stream), "step kernels")) return false;

	common_collect(&c, L->nFluid > 0 ? steps : 0, perLaunch);
	c.t.steps = steps;
	kd->Duration = c.t.loopMs / 1e3;
	if (steps > 0) c.iteration = steps - 1;

	if (!synced) c.shadowStale = 1;
	if (c.shadowDownload || out.vtk || out.stats) {
		const double t0 = (double)clock() / CLOCKS_PER_SEC;
		if (!output_sync(L)) return false;
		c.t.downloadMs = ((double)clock() / CLOCKS_PER_SEC - t0) * 1e3;
	}
	if (out.vtk) VtkWrite(ld, kd, ccd, steps);
	if (out.stats) KernelStatistics(kd, ld, ccd, steps);
	return true;
}

static void list_kernel(LatticeDesc * ld, KernelData * kd, CaseData * cd)
{
	(void)list_steps(kd, cd, ld);
}

ListData * list_data(KernelData * kd)
{
	return (kd && kd->Kernel == list_kernel) ? as_list(kd) : nullptr;
}

// Mass sum_n sum_d f and x momentum sum_n sum_d c_x,d f of the device
// lattice, in FP64 (shifted storage: the stored deviations plus the weights,
// which add 1 per node to the mass and nothing to the momentum).
static bool list_moments(ListData * L, double * mass, double * jx)
{
	Common & c = L->c;
	*mass = 0.0;
	*jx = 0.0;
	if (!list_upload(L)) return false;
	if (L->nFluid == 0) return true;
	const uint32_t nF = (uint32_t)L->nFluid, nC = (uint32_t)L->nCells, nb = (uint32_t)L->momentBlocks;
	double * part = L->dMoments, * sum = L->dMoments + 2 * (size_t)nb;
	const float f16s = (float)(1 << c.args.f16Shift);
	switch (c.args.prec) {
		case PREC_FP32S: k_list_moments<<<nb, 256, 0, c.stream>>>((const float *)L->dPdf, CodecF32{}, nF, nC, part); break;
		case PREC_FP16S:
			k_list_moments<<<nb, 256, 0, c.stream>>>((const uint16_t *)L->dPdf, CodecF16{f16s, 1.0f / f16s}, nF, nC, part);
			break;
		default: k_list_moments<<<nb, 256, 0, c.stream>>>((const double *)L->dPdf, CodecF64{}, nF, nC, part);
	}
	k_sum_pairs<256><<<1, 256, 0, c.stream>>>(part, nb, sum);
	if (!hip_ok(hipGetLastError(), "moment kernels")) return false;
	double h[2];
	if (!hip_ok(hipMemcpyAsync(h, sum, sizeof(h), hipMemcpyDeviceToHost, c.stream), "moment download")) return false;
	if (!hip_ok(hipStreamSynchronize(c.stream), "moment kernels")) return false;
	*mass = h[0] + (c.args.prec == PREC_FP64 ? 0.0 : (double)L->nFluid);
	*jx = h[1];
	return true;
}

// ------------------------------------------------------------ Init/Deinit
static void list_release(ListData * L)
{
	if (!L) return;
	(void)hipFree(L->dGrid);
	(void)hipFree(L->dCoords);
	(void)hipFree(L->dOddOrder);
	(void)hipFree(L->dAdj);
	(void)hipFree(L->dRunMask);
	(void)hipFree(L->dRunBase);
	(void)hipFree(L->dRunAdj);
	(void)hipFree(L->dPdf);
	(void)hipFree(L->dMoments);
	host_free(L->hPdf);
	free(L->hGrid);
	common_free(&L->c);
	delete L;
}


static bool dev_alloc(void ** p, size_t bytes, const char * what)
{
	hipError_t e = hipMalloc(p, bytes ? bytes : 1);
	if (e != hipSuccess) return fail("allocating %s failed: %zu bytes (%s).", what, bytes, hipGetErrorString(e));
	return true;
}

// Run tables for the odd step: start masks per wavefront, a prefix sum of
// the run counts, then the compacted adjacency of every run's first node.
static bool list_build_runs(ListData * L, hipStream_t s)
{
	const uint32_t nF = (uint32_t)L->nFluid;
	L->nWaves = (int)((nF + 63) / 64);
	const unsigned nb = (nF + 255) / 256;
	uint32_t * dCount = nullptr;
	uint32_t total = 0;
	bool ok = false;
	do {
		if (!dev_alloc((void **)&L->dRunMask, sizeof(unsigned long long) * (size_t)L->nWaves, "run masks")) break;
		if (!dev_alloc((void **)&L->dRunBase, sizeof(uint32_t) * (size_t)L->nWaves, "run bases")) break;
		if (!dev_alloc((void **)&dCount, sizeof(uint32_t), "run count")) break;
		k_wave_runs<<<nb, 256, 0, s>>>(L->dAdj, nF, (uint32_t)L->adjStride, (uint32_t)L->nWaves, L->dRunMask, L->dRunBase);
		k_scan_once<<<1, 1024, 0, s>>>(L->dRunBase, (uint32_t)L->nWaves, dCount);
		if (!hip_ok(hipGetLastError(), "run kernels")) break;
		if (!hip_ok(hipMemcpyAsync(&total, dCount, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "run count download")) break;
		if (!hip_ok(hipStreamSynchronize(s), "run kernels")) break;
		L->nRunsWave = (int)total;
		L->runStride = (int)(((total + 255u) / 256u) * 256u);
		if (!dev_alloc((void **)&L->dRunAdj, sizeof(uint32_t) * 18 * (size_t)L->runStride, "run adjacency")) break;
		k_run_compact<<<nb, 256, 0, s>>>(L->dAdj, nF, (uint32_t)L->adjStride, L->dRunMask, L->dRunBase,
		                                 (uint32_t)L->runStride, L->dRunAdj);
		if (!hip_ok(hipGetLastError(), "run compaction")) break;
		ok = true;
	} while (0);
	(void)hipFree(dCount);
	return ok;
}

// y of the first node of every odd-step block
__global__ __launch_bounds__(256) void k_block_y(const uint32_t * __restrict__ coords, uint32_t bs, uint32_t nBlocks,
                                                 uint32_t * __restrict__ y)
{
	const uint32_t b = blockIdx.x * 256 + threadIdx.x;
	if (b < nBlocks) y[b] = coords[3 * (size_t)b * bs + 1];
}

// Launch grid and block order of the odd step (OddOrder).  Band order: the
// device hands launched block b to XCD b % 8; the table gives XCD x the node
// blocks whose first node lies in y band x (of 8), in ascending order, so
// the blocks that share lines of the neighbour planes -- x and y neighbours
// are in the same band -- run in one L2.  Bands hold about equal numbers of
// blocks; the shorter ones are padded with 0xffffffff (no work).
static bool list_build_odd_order(ListData * L, hipStream_t s)
{
	const uint32_t bs = (uint32_t)L->c.args.blockSize;
	const uint32_t nb = (uint32_t)((L->nFluid + bs - 1) / bs);
	if (L->c.args.oddOrder != 2 || nb == 0) {
		const uint32_t g = L->c.args.oddOrder == 1 ? 8 * kXcdGroup : 1;
		L->oddGrid = (nb + g - 1) / g * g;
		return true;
	}
	uint32_t * dY = nullptr;
	if (!dev_alloc((void **)&dY, sizeof(uint32_t) * nb, "block rows")) return false;
	std::vector<uint32_t> y(nb);
	k_block_y<<<(nb + 255) / 256, 256, 0, s>>>(L->dCoords, bs, nb, dY);
	bool ok = hip_ok(hipGetLastError(), "block rows") &&
	          hip_ok(hipMemcpyAsync(y.data(), dY, sizeof(uint32_t) * nb, hipMemcpyDeviceToHost, s), "block rows") &&
	          hip_ok(hipStreamSynchronize(s), "block rows");
	(void)hipFree(dY);
	if (!ok) return false;
	std::vector<uint32_t> band[8];
	const uint64_t Y = (uint64_t)L->dims[1];
	for (uint32_t b = 0; b < nb; ++b) band[(uint64_t)y[b] * 8 / Y].push_back(b);
	size_t longest = 0;
	for (const auto & v : band) longest = v.size() > longest ? v.size() : longest;
	std::vector<uint32_t> table(8 * longest, 0xffffffffu);
	for (uint32_t x = 0; x < 8; ++x)
		for (size_t j = 0; j < band[x].size(); ++j) table[8 * j + x] = band[x][j];
	L->oddGrid = (uint32_t)table.size();
	if (!dev_alloc((void **)&L->dOddOrder, sizeof(uint32_t) * table.size(), "odd-step block order")) return false;
	return hip_ok(hipMemcpyAsync(L->dOddOrder, table.data(), sizeof(uint32_t) * table.size(), hipMemcpyHostToDevice, s),
	              "odd-step block order") &&
	       hip_ok(hipStreamSynchronize(s), "odd-step block order");
}

static bool list_build(ListData * L, const LatticeDesc * ld)
{
	const int X = L->dims[0], Y = L->dims[1], Z = L->dims[2];
	const uint32_t N = (uint32_t)X * Y * Z;
	const uint32_t nBlocks = (N + kNumBlock - 1) / kNumBlock;
	int * dMask = nullptr;
	uint32_t * dSums = nullptr;
	uint32_t hTotal[2] = {0, 0};
	bool ok = false;

	do {
		if (!dev_alloc((void **)&dMask, sizeof(int) * (size_t)N, "geometry mask")) break;
		if (!dev_alloc((void **)&dSums, sizeof(uint32_t) * ((size_t)nBlocks + 2), "scan buffer")) break;
		if (!dev_alloc((void **)&L->dGrid, sizeof(uint32_t) * (size_t)N, "grid")) break;
		hipStream_t s = L->c.stream;
		if (!hip_ok(hipMemcpyAsync(dMask, ld->Lattice, sizeof(int) * (size_t)N, hipMemcpyHostToDevice, s), "mask upload")) break;

		// count first: coords and adjacency are sized by the count, which
		// must agree with LatticeDesc.nFluid before anything is written
		NumGeom g = {X, Y, Z, L->blk[0], L->blk[1], L->blk[2]};
		(void)hipGetLastError();
		k_num_count<<<nBlocks, kNumBlock, 0, s>>>(dMask, g, N, dSums);
		k_scan_once<<<1, 1024, 0, s>>>(dSums, nBlocks, dSums + nBlocks);
		if (!hip_ok(hipGetLastError(), "numbering kernels")) break;
		if (!hip_ok(hipMemcpyAsync(&hTotal[0], dSums + nBlocks, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "count download")) break;
		if (!hip_ok(hipStreamSynchronize(s), "numbering kernels")) break;
		if ((int)hTotal[0] != L->nFluid) {
			fail("list build: %u non-solid cells in the mask, LatticeDesc.nFluid = %d.", hTotal[0], L->nFluid);
			break;
		}
		if (!dev_alloc((void **)&L->dCoords, sizeof(uint32_t) * 3 * (size_t)(L->nFluid > 0 ? L->nFluid : 1), "coords")) break;
		if (!dev_alloc((void **)&L->dAdj, sizeof(uint32_t) * 18 * (size_t)L->adjStride, "adjacency")) break;
		if (!hip_ok(hipMemsetAsync(L->dAdj, 0, sizeof(uint32_t) * 18 * (size_t)L->adjStride, s), "hipMemset")) break;
		k_num_assign<<<nBlocks, kNumBlock, 0, s>>>(dMask, g, N, dSums, L->dGrid, L->dCoords);
		if (!hip_ok(hipGetLastError(), "numbering kernels")) break;
		if (L->nFluid > 0) {
			AdjGeom ag = {X, Y, Z, L->periodic[0], L->periodic[1], L->periodic[2],
			              (uint32_t)L->nFluid, (uint32_t)L->nCells, (uint32_t)L->adjStride};
			const unsigned nb = (unsigned)((L->nFluid + 255) / 256);
			k_adjacency<<<nb, 256, 0, s>>>(dMask, L->dGrid, L->dCoords, ag, L->dAdj);
			if (!hip_ok(hipMemsetAsync(dSums, 0, sizeof(uint32_t), s), "hipMemset")) break;
			k_run_breaks<<<nb, 256, 0, s>>>(L->dAdj, (uint32_t)L->nFluid, (uint32_t)L->adjStride, dSums);
			if (!hip_ok(hipGetLastError(), "adjacency kernels")) break;
			if (!hip_ok(hipMemcpyAsync(&hTotal[1], dSums, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "run count download")) break;
		}
		if (L->nFluid > 0 && !list_build_runs(L, s)) break;
		if (!list_build_odd_order(L, s)) break;
		L->hGrid = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)N);
		if (!L->hGrid) { fail("allocating host grid failed: %zu bytes.", sizeof(uint32_t) * (size_t)N); break; }
		if (!hip_ok(hipMemcpyAsync(L->hGrid, L->dGrid, sizeof(uint32_t) * (size_t)N, hipMemcpyDeviceToHost, s), "grid download")) break;
		if (!hip_ok(hipStreamSynchronize(s), "adjacency kernels")) break;
		L->nRuns = L->nFluid > 0 ? (int)hTotal[1] + 1 : 0;
		ok = true;
	} while (0);

	(void)hipFree(dMask);
	(void)hipFree(dSums);
	return ok;
}

static bool list_download_kd(KernelData * kd) { return list_download(as_list(kd)); }

Common * list_common(KernelData * kd)
{
	return (kd && kd->Kernel == list_kernel) ? &as_list(kd)->c : nullptr;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

void D3Q19ListAaPvInit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params)
{
	*kernelData = nullptr;
	ListData * L = new (std::nothrow) ListData();
	if (!L) { fail("allocating kernel data failed."); return; }
	memset(&L->kd, 0, sizeof(L->kd));
	if (!common_init(&L->c, params, true)) { list_release(L); return; }

	for (int i = 0; i < 3; ++i) {
		L->dims[i] = ld->Dims[i];
		L->blk[i] = L->c.args.blk[i] ? L->c.args.blk[i] : ld->Dims[i];
	}
	L->periodic[0] = ld->PeriodicX; L->periodic[1] = ld->PeriodicY; L->periodic[2] = ld->PeriodicZ;
	L->nFluid = ld->nFluid;

	const uint64_t nTotal = (uint64_t)ld->Dims[0] * ld->Dims[1] * ld->Dims[2];
	if (ld->Dims[0] <= 0 || ld->Dims[1] <= 0 || ld->Dims[2] <= 0 || 19ull * nTotal > 0xffffffffull) {
		fail("number of PDFs exceed 2^32 or empty dimensions.");   // src/Main.c:399-406
		list_release(L);
		return;
	}
	const long nCells = pad_cells(L->nFluid, (int)sizeof(PdfT), L->c.args.havePad ? &L->c.args.pad : nullptr);
	if (nCells < 0 || 19ull * (uint64_t)nCells > 0xffffffffull) {
		fail("padded node count %ld does not fit the 32-bit adjacency.", nCells);
		list_release(L);
		return;
	}
	L->nCells = (int)nCells;
	L->adjStride = ((L->nFluid + 255) / 256) * 256;

	if (!L->c.args.quiet) {
		printf("# blocking               x: %3d y: %3d z: %3d\n", L->blk[0], L->blk[1], L->blk[2]);
		printf("# lattice size:          %e MiB\n", nCells * 8.0 * 19 / 1024.0 / 1024.0);
		printf("# fluid lattice size:    %e MiB\n", L->nFluid * 8.0 * 19 / 1024.0 / 1024.0);
		printf("# index size:            %e MiB\n", L->nFluid * 4.0 * 18 / 1024.0 / 1024.0);
	}

	if (!list_build(L, ld)) { list_release(L); return; }

	// zero bits are 0.0 in every storage type: for shifted storage they mean
	// f = w, for FP64 f = 0 as the reference initialises
	// (src/BenchKernelD3Q19ListAaPvCommon.c:583-585)
	const size_t pdfBytes = elem_bytes(L->c.args.prec) * 19 * (size_t)L->nCells;
	if (!dev_alloc((void **)&L->dPdf, pdfBytes, "device lattice")) { list_release(L); return; }
	if (!hip_ok(hipMemsetAsync(L->dPdf, 0, pdfBytes, L->c.stream), "hipMemset")) { list_release(L); return; }
	L->momentBlocks = (L->nFluid + 255) / 256;
	if (!dev_alloc((void **)&L->dMoments, sizeof(double) * (2 * (size_t)L->momentBlocks + 2), "moment scratch")) { list_release(L); return; }
	if (!L->c.args.lazyShadow) {
		L->hPdf = host_alloc(pdfBytes);
		if (!L->hPdf) { fail("allocating host lattice failed: %zu bytes.", pdfBytes); list_release(L); return; }
		memset(L->hPdf, 0, pdfBytes);
	}
	if (!hip_ok(hipStreamSynchronize(L->c.stream), "init")) { list_release(L); return; }

	const double even = list_even_bytes(L), odd = list_odd_bytes(L);
	if (!L->c.args.quiet) {
		printf("# total fluid nodes: %d   consecutive blocks: %d   runs cut at wavefronts: %d\n",
		       L->nFluid, L->nRuns, L->nRunsWave);
		printf("# loop balance:          %.2f B/FLUP  even: %.2f B/FLUP  odd %.2f B/FLUP\n", (even + odd) / 2, even, odd);
	}

	KernelData * kd = &L->kd;
	kd->Pdfs[0] = (PdfT *)L->hPdf;     // storage type given by -prec (double by default)
	kd->PdfsActive = (PdfT *)L->hPdf;
	for (int i = 0; i < 3; ++i) { kd->Dims[i] = ld->Dims[i]; kd->GlobalDims[i] = ld->Dims[i]; kd->Offsets[i] = 0; }
	kd->LoopBalance = (even + odd) / 2;
	kd->GetNode = list_get_node;
	kd->SetNode = list_set_node;
	kd->BoundaryConditionsGetPdf = list_bc_get;
	kd->BoundaryConditionsSetPdf = list_bc_set;
	kd->Kernel = list_kernel;
	L->c.download = list_download_kd;
	*kernelData = kd;
}

void D3Q19ListAaPvDeinit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData)
{
	(void)ld;
	if (kernelData && *kernelData) {
		list_release(as_list(*kernelData));
		*kernelData = nullptr;
	}
}

int LbmHipListGetInfo(const KernelData * kd, LbmHipListInfo * out)
{
	if (!kd || !out || kd->Kernel != list_kernel) { fail("LbmHipListGetInfo: not a list-aa-pv-soa kernel."); return -1; }
	const ListData * L = as_list(kd);
	out->nFluid = L->nFluid;
	out->nCells = L->nCells;
	out->adjStride = L->adjStride;
	out->nRuns = L->nRuns;
	for (int i = 0; i < 3; ++i) { out->dims[i] = L->dims[i]; out->periodic[i] = L->periodic[i]; }
	out->bytesPerFlupEven = list_even_bytes(L);
	out->prec = L->c.args.prec;
	out->bytesPerFlupOdd = list_odd_bytes(L);
	out->nRunsWave = L->nRunsWave;
	out->adjRle = L->c.args.adjRle;
	return 0;
}

}  // extern "C"

namespace lbm {

// Either kernel: moment sums, a chunk of steps, and the final download.
static bool any_moments(KernelData * kd, double * m, double * j)
{
	return list_data(kd) ? list_moments(list_data(kd), m, j) : aa_moments(kd, m, j);
}

static bool any_steps(KernelData * kd, const CaseData * cd)
{
	return list_data(kd) ? list_steps(kd, cd, nullptr) : aa_steps(kd, cd);
}

static bool any_download(KernelData * kd)
{
	return list_data(kd) ? list_download(list_data(kd)) : aa_download_shadow(kd);
}

// Kernel() chunks of checkEvery steps (rounded up to even) until the
// relative change of the x momentum between checks is at most tol, or
// maxSteps steps ran.
static int run_steady(KernelData * kd, CaseData * cd, int checkEvery, double tol, int maxSteps, LbmHipSteady * out,
                      const char * who)
{
	if (!cd || checkEvery <= 0 || maxSteps <= 0 || !(tol >= 0.0)) {
		fail("%s: needs a case, checkEvery > 0, maxSteps > 0 and tol >= 0.", who);
		return -1;
	}
	Common & c = *(list_data(kd) ? list_common(kd) : aa_common(kd));
	const int every = checkEvery + (checkEvery & 1);
	const int saveDownload = c.shadowDownload;
	c.shadowDownload = 0;
	LbmHipSteady r = {};
	r.relChange = -1.0;
	double prev = 0.0, seconds = 0.0;
	bool ok = true, havePrev = false;
	while (r.steps < maxSteps) {
		CaseData chunk = *cd;
		chunk.MaxIterations = every < maxSteps - r.steps ? every : maxSteps - r.steps;
		if (!any_steps(kd, &chunk)) { ok = false; break; }
		r.steps += chunk.MaxIterations + (chunk.MaxIterations & 1);
		seconds += kd->Duration;
		if (!any_moments(kd, &r.mass, &r.momentumX)) { ok = false; break; }
		if (havePrev) {
			const double d = fabs(r.momentumX - prev), a = fabs(r.momentumX);
			r.relChange = a > 0.0 ? d / a : (d > 0.0 ? INFINITY : 0.0);
			if (r.relChange <= tol) { r.converged = 1; break; }
		}
		prev = r.momentumX;
		havePrev = true;
	}
	c.shadowDownload = saveDownload;
	kd->Duration = seconds;
	r.seconds = seconds;
	if (ok && saveDownload && !any_download(kd)) ok = false;
	if (out) *out = r;
	return ok ? 0 : -1;
}

}  // namespace lbm

extern "C" {

int LbmHipListMoments(KernelData * kd, double * mass, double * momentumX)
{
	if (!kd || kd->Kernel != list_kernel) { fail("LbmHipListMoments: not a list-aa-pv-soa kernel."); return -1; }
	return LbmHipMoments(kd, mass, momentumX);
}

int LbmHipListRunSteady(KernelData * kd, CaseData * cd, int checkEvery, double tol, int maxSteps, LbmHipSteady * out)
{
	if (!kd || kd->Kernel != list_kernel) { fail("LbmHipListRunSteady: not a list-aa-pv-soa kernel."); return -1; }
	return run_steady(kd, cd, checkEvery, tol, maxSteps, out, "LbmHipListRunSteady");
}

int LbmHipMoments(KernelData * kd, double * mass, double * momentumX)
{
	if (!list_data(kd) && !aa_common(kd)) { fail("LbmHipMoments: not a kernel of this library."); return -1; }
	double m = 0.0, j = 0.0;
	if (!any_moments(kd, &m, &j)) return -1;
	if (mass) *mass = m;
	if (momentumX) *momentumX = j;
	return 0;
}

int LbmHipRunSteady(KernelData * kd, CaseData * cd, int checkEvery, double tol, int maxSteps, LbmHipSteady * out)
{
	if (!list_data(kd) && !aa_common(kd)) { fail("LbmHipRunSteady: not a kernel of this library."); return -1; }
	return run_steady(kd, cd, checkEvery, tol, maxSteps, out, "LbmHipRunSteady");
}

int LbmHipListCopyBuild(const KernelData * kd, uint32_t * adj, uint32_t * grid, uint32_t * coords)
{
	if (!kd || kd->Kernel != list_kernel) { fail("LbmHipListCopyBuild: not a list-aa-pv-soa kernel."); return -1; }
	const ListData * L = as_list(kd);
	const size_t N = (size_t)L->dims[0] * L->dims[1] * L->dims[2];
	if (grid && !hip_ok(hipMemcpy(grid, L->dGrid, sizeof(uint32_t) * N, hipMemcpyDeviceToHost), "grid download")) return -1;
	if (coords && L->nFluid > 0 &&
	    !hip_ok(hipMemcpy(coords, L->dCoords, sizeof(uint32_t) * 3 * (size_t)L->nFluid, hipMemcpyDeviceToHost), "coords download")) return -1;
	if (adj && L->nFluid > 0) {
		uint32_t * t = (uint32_t *)malloc(sizeof(uint32_t) * 18 * (size_t)L->adjStride);
		if (!t) { fail("allocating %zu bytes failed.", sizeof(uint32_t) * 18 * (size_t)L->adjStride); return -1; }
		if (!hip_ok(hipMemcpy(t, L->dAdj, sizeof(uint32_t) * 18 * (size_t)L->adjStride, hipMemcpyDeviceToHost), "adjacency download")) { free(t); return -1; }
		for (size_t n = 0; n < (size_t)L->nFluid; ++n)
			for (int d = 0; d < 18; ++d) adj[n * 18 + d] = t[(size_t)d * L->adjStride + n];
		free(t);
	}
	return 0;
}

}  // extern "C"

// ------------------------------------------- list build without allocation
// LbmHipListBuildNoAlloc runs the list build of D3Q19ListAaPvInit_PushSoAHip
// (k_num_count, k_scan_once, k_num_assign, k_adjacency) into device buffers
// the caller owns.  Its only scratch -- the per-block fluid counts of the
// numbering and their total -- lives in the file-scope array below instead of
// a hipMalloc'ed buffer.  Each GPU holds its own copy of the array until the
// process ends; all calls on one GPU share it and leave the last call's
// contents, so calls that overlap on one GPU race.  Host code reaches the
// array only through the device address hipGetSymbolAddress returns for the
// current device (g_scratchDev), never through its name.
namespace lbm {

constexpr size_t kScratchBytes = size_t(1) << 20;                                      // 1 MiB
constexpr uint32_t kScratchMaxBlocks = (uint32_t)(kScratchBytes / sizeof(uint32_t)) - 1;   // + the total

static __device__ __align__(256) uint32_t scratch[kScratchBytes / sizeof(uint32_t)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratchMutex;
static uint32_t * g_scratchDev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
static uint32_t * scratch_address()
{
	int dev = -1;
	if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
	if (dev < 0 || dev >= kMaxDevices) {
		fail("LbmHipListBuildNoAlloc: device %d outside the %d devices this library tracks.", dev, kMaxDevices);
		return nullptr;
	}
	std::lock_guard<std::mutex> lock(g_scratchMutex);
	if (!g_scratchDev[dev]) {
		void * p = nullptr;
		if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return nullptr;
		g_scratchDev[dev] = static_cast<uint32_t *>(p);
	}
	return g_scratchDev[dev];
}

}  // namespace lbm

extern "C" int LbmHipListBuildNoAlloc(const int * dMask, const int dims[3], const int periodic[3], const int blk[3],
                                      int nCells, int adjStride, int capacity, uint32_t * dGrid, uint32_t * dCoords,
                                      uint32_t * dAdj, int * nFluid)
{
	if (nFluid) *nFluid = 0;
	if (!dMask || !dims || !periodic || !dGrid || !dCoords || !dAdj || !nFluid) {
		fail("LbmHipListBuildNoAlloc: null argument.");
		return -1;
	}
	const int X = dims[0], Y = dims[1], Z = dims[2];
	if (X <= 0 || Y <= 0 || Z <= 0 || 19ull * (uint64_t)X * Y * Z > 0xffffffffull) {
		fail("number of PDFs exceed 2^32 or empty dimensions.");   // src/Main.c:399-406
		return -1;
	}
	if (nCells < 0 || adjStride < 0 || capacity < 0 || 19ull * (uint64_t)nCells > 0xffffffffull) {
		fail("LbmHipListBuildNoAlloc: nCells %d, adjStride %d, capacity %d out of range.", nCells, adjStride, capacity);
		return -1;
	}
	int b[3];
	for (int i = 0; i < 3; ++i) {
		const int v = blk ? blk[i] : 0;
		if (v < 0) { fail("LbmHipListBuildNoAlloc: negative block size %d.", v); return -1; }
		b[i] = v ? v : dims[i];
	}
	const uint32_t N = (uint32_t)X * Y * Z;
	const uint32_t nBlocks = (N + kNumBlock - 1) / kNumBlock;
	if (nBlocks > kScratchMaxBlocks) {
		fail("LbmHipListBuildNoAlloc: %u cells exceed the scratch limit of %u cells (%zu bytes).", N,
		     kScratchMaxBlocks * (uint32_t)kNumBlock, kScratchBytes);
		return -1;
	}
	uint32_t * s = scratch_address();
	if (!s) return -1;

	hipStream_t st = nullptr;
	(void)hipGetLastError();
	const NumGeom g = {X, Y, Z, b[0], b[1], b[2]};
	k_num_count<<<nBlocks, kNumBlock, 0, st>>>(dMask, g, N, s);
	k_scan_once<<<1, 1024, 0, st>>>(s, nBlocks, s + nBlocks);
	if (!hip_ok(hipGetLastError(), "numbering kernels")) return -1;
	uint32_t total = 0;
	if (!hip_ok(hipMemcpyAsync(&total, s + nBlocks, sizeof(total), hipMemcpyDeviceToHost, st), "count download")) return -1;
	if (!hip_ok(hipStreamSynchronize(st), "numbering kernels")) return -1;
	if (total > (uint32_t)capacity || total > (uint32_t)nCells || total > (uint32_t)adjStride) {
		fail("LbmHipListBuildNoAlloc: %u fluid cells, room for %d nodes (nCells %d, adjStride %d).", total, capacity,
		     nCells, adjStride);
		return -1;
	}
	k_num_assign<<<nBlocks, kNumBlock, 0, st>>>(dMask, g, N, s, dGrid, dCoords);
	if (total > 0) {
		const AdjGeom ag = {X, Y, Z, periodic[0], periodic[1], periodic[2], total, (uint32_t)nCells,
		                    (uint32_t)adjStride};
		k_adjacency<<<(total + 255) / 256, 256, 0, st>>>(dMask, dGrid, dCoords, ag, dAdj);
	}
	if (!hip_ok(hipGetLastError(), "adjacency kernels")) return -1;
	if (!hip_ok(hipStreamSynchronize(st), "adjacency kernels")) return -1;
	*nFluid = (int)total;
	return 0;
}

// ------------------------------------------------------ device footprint
// Upper bound of the device memory D3Q19ListAaPvInit_PushSoAHip takes for
// a lattice, from the same sizes Init derives (padding, strides), without
// touching the device: every allocation Init makes, the build's temporary
// buffers included, the run-compressed adjacency at its largest (one run
// per node) and each allocation rounded up to the allocator's 2 MiB.
extern "C" int LbmHipListDeviceBytes(const LatticeDesc * ld, const Parameters * params, uint64_t * bytes)
{
	if (!ld || !bytes) { fail("LbmHipListDeviceBytes: null argument."); return -1; }
	*bytes = 0;
	KernelArgs a;
	if (!parse_kernel_args(params, &a, true)) return -1;
	if (ld->Dims[0] <= 0 || ld->Dims[1] <= 0 || ld->Dims[2] <= 0 || ld->nFluid < 0) {
		fail("LbmHipListDeviceBytes: empty dimensions.");
		return -1;
	}
	const uint64_t N = (uint64_t)ld->Dims[0] * ld->Dims[1] * ld->Dims[2];
	const uint64_t nF = (uint64_t)ld->nFluid;
	// what Init refuses, refused here too
	if (19ull * N > 0xffffffffull) { fail("number of PDFs exceed 2^32 or empty dimensions."); return -1; }
	const long nCells = pad_cells(ld->nFluid, (int)sizeof(PdfT), a.havePad ? &a.pad : nullptr);
	if (nCells < 0) { fail("LbmHipListDeviceBytes: padding failed."); return -1; }
	if (19ull * (uint64_t)nCells > 0xffffffffull) {
		fail("padded node count %ld does not fit the 32-bit adjacency.", nCells);
		return -1;
	}
	const uint64_t adjStride = (nF + 255) / 256 * 256, nWaves = (nF + 63) / 64;
	const uint64_t nBlocks = (N + kNumBlock - 1) / kNumBlock, oddBlocks = (nF + a.blockSize - 1) / a.blockSize;
	const uint64_t sizes[] = {
		elem_bytes(a.prec) * 19 * (uint64_t)nCells,       // PDFs
		4 * 18 * adjStride,                                // adjacency
		4 * N,                                             // grid
		12 * (nF > 0 ? nF : 1),                            // coords
		8 * (2 * ((nF + 255) / 256) + 2),                  // moment sums
		8 * nWaves, 4 * nWaves, 4 * 18 * adjStride,        // run masks, bases, run adjacency (at most one run per node)
		4 * 8 * (oddBlocks + 8),                           // odd-step block order
		4 * N, 4 * (nBlocks + 2), 4, 4 * (oddBlocks + 1),  // build temporaries: mask, scan, run count, block rows
	};
	const uint64_t page = 2ull << 20;
	uint64_t total = 0;
	for (uint64_t s : sizes) total += (s + page - 1) / page * page;
	*bytes = total;
	return 0;
}
