// lbm_aa.hip -- aa-soa on MI355X: full-array AA pattern with one ghost
// layer, behind the reference's kernel-plugin ABI.
//
// The reference (src/BenchKernelD3Q19AaCommon.c:309-637,
// src/BenchKernelD3Q19Aa.c:40-548) streams through a ghost-padded array and,
// after each step, copies PDFs along precomputed (src, dst) index pairs to
// realise half-way bounce-back and periodic wrap-around.  Here the same
// data movement is resolved while the odd step loads and stores: a per-cell
// info word carries one bounce-back bit per direction, and cells on the
// domain boundary compute the periodic image of their neighbour.  Every
// fluid value ends up in the slot the reference's fix-up copies would have
// put it, so node views are bit-identical to the reference.  Pairs whose
// target lies outside a non-periodic face have no fix-up in the reference;
// the value stays in the ghost cell and is read back from there by the same
// node two steps later.  The ghost layer exists for exactly that case and
// starts at f = 0 (the reference leaves it unset; fresh pages read 0.0).
//
// Device layout: PDF(d, x, y, z) at
//   d * plane + ((x + 1) * gY + (y + 1)) * zStride + (z + zOff)
// gY = Y + 2, plane = (X + 2) * gY * zStride.  Rows with z ghost cells:
// zOff = 128 B of stored values (16 FP64, 32 FP32 or 64 FP16 values; the
// ghost z = -1 sits at zOff - 1) and zStride a multiple of zOff, so that
// interior z rows start on a 128-B line and a wavefront's 64 lanes cover
// whole lines (four, two or one).  Compact rows (zOff 0, zStride Z rounded
// up to whole 128-B lines), for FP64 and FP32 storage when no fluid cell can
// reach a z ghost cell (z periodic, or no fluid cell on either z face): the
// rows of an x slab are then contiguous when Z fills whole lines, which
// streams 14 % faster than rows with a 256-B gap at 256^3.
//
// Storage: FP64 values, or with -prec fp32s / fp16s the deviations f - w_d
// through the list kernel's codecs (the narrow even step takes 2 / 4 cells
// of a z row per thread, the FP32 odd step 2).  Run-time VTK / statistics
// output follows the reference's call points, including the views after an
// even step.
#include "lbm_list.h"      // storage codecs, elem_bytes, kW

// The program's output functions (the reference's Vtk.o / Kernel.o, or
// liblbmhost.so); weak, so the library loads without them.
extern "C" {
void VtkWrite(LatticeDesc * ld, KernelData * kd, CaseData * cd, int iteration) __attribute__((weak));
void KernelStatistics(KernelData * kd, LatticeDesc * ld, CaseData * cd, int iteration) __attribute__((weak));
}

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <new>

namespace lbm {

constexpr uint32_t kInfoFluid = 1u << 31;
constexpr uint32_t kInfoBoundary = 1u << 30;

struct AaGeom {
	int X, Y, Z;
	int px, py, pz;
	int gY;
	int zStride;
	int zOff;                // 128 B of stored values
	int zWords;              // 64-bit words of the fluid bit mask per z row
	uint64_t plane;
};

struct AaData {
	KernelData kd;
	Common c;
	AaGeom g = {};
	int nFluid = 0;
	int nBounce = 0;
	uint32_t * dInfo = nullptr;   // [x][y][z], z fastest
	unsigned long long * dFluid = nullptr;   // fluid bit per cell, [x][y][z / 64] words (even step)
	void * dPdf = nullptr;        // nPdf values of the storage type (-prec)
	void * hPdf = nullptr;
	uint32_t * hInfo = nullptr;
	size_t nPdf = 0;
	double * dMoments = nullptr;  // per-block pairs, then the total (allocated at first use)
	uint32_t momentBlocks = 0;
	// run-time output after an even step: GetNode serves the reference's
	// view at even parity (pull view, force added) from the host copy
	bool pullView = false;
	double viewFt[19] = {};
};

static AaData * as_aa(KernelData * kd) { return reinterpret_cast<AaData *>(kd); }

__host__ __device__ __forceinline__ uint64_t aa_slot(const AaGeom & g, int d, int x, int y, int z)
{
	return (uint64_t)d * g.plane + ((uint64_t)(x + 1) * g.gY + (uint64_t)(y + 1)) * g.zStride + (uint64_t)(z + g.zOff);
}

// Neighbour (x + dx, y + dy, z + dz) after the reference's periodic rule:
// when a coordinate leaves the box through a periodic face, every
// out-of-range coordinate is wrapped (src/BenchKernelD3Q19AaCommon.c:522-567);
// otherwise out-of-range coordinates stay in the ghost layer.
__host__ __device__ __forceinline__ void aa_resolve(const AaGeom & g, int & x, int & y, int & z)
{
	const bool ox = x < 0 || x >= g.X, oy = y < 0 || y >= g.Y, oz = z < 0 || z >= g.Z;
	if ((ox && g.px) || (oy && g.py) || (oz && g.pz)) {
		x = x < 0 ? g.X - 1 : (x >= g.X ? 0 : x);
		y = y < 0 ? g.Y - 1 : (y >= g.Y ? 0 : y);
		z = z < 0 ? g.Z - 1 : (z >= g.Z ? 0 : z);
	}
}

// Info word per cell: fluid bit, boundary bit (a neighbour lies outside the
// box), and bit d set when the push target of direction d is a solid cell
// (in the box or as periodic image) -- the pairs for which the reference's
// fix-up list holds a bounce-back entry.
__global__ __launch_bounds__(256) void k_aa_info(const int * __restrict__ mask, AaGeom g, uint32_t * __restrict__ info,
                                                 uint32_t * __restrict__ nBounce)
{
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	const uint32_t n = (uint32_t)g.X * g.Y * g.Z;
	int bounces = 0;
	if (i < n) {
		const int z = (int)(i % g.Z), y = (int)((i / g.Z) % g.Y), x = (int)(i / ((uint32_t)g.Z * g.Y));
		uint32_t w = 0;
		if (mask[((size_t)z * g.Y + y) * g.X + x] != LBM_CELL_OBSTACLE) {
			w |= kInfoFluid;
			if (x == 0 || y == 0 || z == 0 || x == g.X - 1 || y == g.Y - 1 || z == g.Z - 1) w |= kInfoBoundary;
			for (int d = 0; d < 18; ++d) {
				int nx = x + cx(d), ny = y + cy(d), nz = z + cz(d);
				aa_resolve(g, nx, ny, nz);
				const bool inside = nx >= 0 && nx < g.X && ny >= 0 && ny < g.Y && nz >= 0 && nz < g.Z;
				if (inside && mask[((size_t)nz * g.Y + ny) * g.X + nx] == LBM_CELL_OBSTACLE) {
					w |= 1u << d;
					++bounces;
				}
			}
		}
		info[i] = w;
	}
	// total bounce-back pairs (reported like the reference's fix-up count)
	for (int off = 32; off > 0; off >>= 1) bounces += __shfl_down(bounces, off);
	if (__lane_id() == 0 && bounces) atomicAdd(nBounce, (uint32_t)bounces);
}

// Fluid bit of every cell, 64 cells of a z row per word: what the even step
// reads instead of the 4-byte info word.
__global__ __launch_bounds__(256) void k_aa_fluid_bits(const uint32_t * __restrict__ info, AaGeom g,
                                                       unsigned long long * __restrict__ bits)
{
	const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= (uint64_t)g.X * g.Y * g.zWords) return;
	const uint64_t row = i / (uint64_t)g.zWords;
	const int z0 = (int)(i % (uint64_t)g.zWords) * 64;
	unsigned long long w = 0;
	for (int j = 0; j < 64 && z0 + j < g.Z; ++j)
		if (info[row * (uint64_t)g.Z + z0 + j] & kInfoFluid) w |= 1ull << j;
	bits[i] = w;
}

__device__ __forceinline__ unsigned long long aa_fluid_word(const unsigned long long * bits, const AaGeom & g, int x, int y,
                                                            int z)
{
	return bits[((size_t)x * g.Y + y) * g.zWords + (z >> 6)];
}

// Cell of this thread: z along threadIdx.x, blockDim.y rows of y per block.
__device__ __forceinline__ bool aa_cell(const AaGeom & g, int & x, int & y, int & z)
{
	z = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
	x = (int)blockIdx.z;
	return z < g.Z && y < g.Y;
}

// Even step: local load, collision, store of direction d into slot inv(d)
// of the same cell (src/BenchKernelD3Q19Aa.c:146-290).
// NT: non-temporal loads and stores (each value is touched once per step).
// Solid cells of the box load their slots and store them back unchanged, so
// that every wavefront writes whole lines (a non-temporal store of a partly
// written line costs far more than the bytes saved; no fluid cell ever reads
// a solid cell's slots).
template <class Codec, bool FORCE_LOAD, bool NT>
__global__ __launch_bounds__(512) void k_aa_even(typename Codec::S * pdf, Codec cod,
                                                 const unsigned long long * __restrict__ fluidBits, AaGeom g,
                                                 TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft)
{
	using R = typename Codec::R;
	int x, y, z;
	if (!aa_cell(g, x, y, z)) return;
	const bool fluid = (aa_fluid_word(fluidBits, g, x, y, z) >> (z & 63)) & 1ull;
	const uint64_t base = aa_slot(g, 0, x, y, z);
	R f[19];
#pragma unroll
	for (int d = 0; d < 19; ++d) f[d] = cod.template load<NT>(&pdf[base + (uint64_t)d * g.plane]);
	if (fluid) {
		if (FORCE_LOAD) {
#pragma unroll
			for (int d = 0; d < 19; ++d) f[d] = f[d] + ft.f[d];
		}
		trt_collide(f, k);
	}
	// slot inv(d) takes f_d of a fluid cell, its own value for a solid one
#pragma unroll
	for (int d = 0; d < 19; ++d)
		cod.template store<NT>(&pdf[base + (uint64_t)inv_dir(d) * g.plane], fluid ? f[d] : f[inv_dir(d)]);
}

// Even step over V consecutive cells z0 .. z0+V-1 of a z row, for the
// narrow storage types: one 8-byte access per direction carries V values
// (V = 2 for FP32, 4 for FP16).  zOff, zStride and plane are multiples of V,
// so the accesses are aligned; a group that reaches past the row end stays
// in the row's padding (zStride is a multiple of V).  Cells of the group that are solid or past
// the row end (ghost z = Z, padding) get their values written back
// unchanged: the slot that takes direction d of a fluid cell, inv(d), takes
// the cell's own value inv(d) (a decode/encode round trip is exact).  Each
// fluid cell's arithmetic is that of k_aa_even.
template <class Codec, int V, bool FORCE_LOAD, bool NT>
__global__ __launch_bounds__(512) void k_aa_even_v(typename Codec::S * pdf, Codec cod,
                                                   const unsigned long long * __restrict__ fluidBits, AaGeom g,
                                                   TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft)
{
	using R = typename Codec::R;
	using VT = VecS<typename Codec::S, V>;
	const int z0 = (int)(blockIdx.x * blockDim.x + threadIdx.x) * V;
	const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
	const int x = (int)blockIdx.z;
	if (z0 >= g.Z || y >= g.Y) return;
	// z0 .. z0+V-1 lie in one word of the mask (V divides 64)
	const unsigned long long word = aa_fluid_word(fluidBits, g, x, y, z0) >> (z0 & 63);
	bool fluid[V];
	bool any = false;
#pragma unroll
	for (int j = 0; j < V; ++j) {
		fluid[j] = z0 + j < g.Z && ((word >> j) & 1ull);
		any = any || fluid[j];
	}
	if (!any) return;
	const uint64_t base = aa_slot(g, 0, x, y, z0);
	R f[V][19];
#pragma unroll
	for (int d = 0; d < 19; ++d) {
		const VT v = ld<NT>(reinterpret_cast<const VT *>(&pdf[base + (uint64_t)d * g.plane]));
#pragma unroll
		for (int j = 0; j < V; ++j) f[j][d] = cod.dec(v[j]);
	}
#pragma unroll
	for (int j = 0; j < V; ++j) {
		if (fluid[j]) {
			if (FORCE_LOAD) {
#pragma unroll
				for (int d = 0; d < 19; ++d) f[j][d] = f[j][d] + ft.f[d];
			}
			trt_collide(f[j], k);
		}
	}
#pragma unroll
	for (int d = 0; d < 19; ++d) {
		VT v;
#pragma unroll
		for (int j = 0; j < V; ++j) v[j] = cod.enc(fluid[j] ? f[j][d] : f[j][inv_dir(d)]);
		st<NT>(reinterpret_cast<VT *>(&pdf[base + (uint64_t)inv_dir(d) * g.plane]), v);
	}
}

// Odd step of one fluid cell x with info word w: direction d arrives from
// the pull neighbour x - c_d, slot inv(d)
// (src/BenchKernelD3Q19Aa.c:392) -- or from the cell's own slot d when that
// neighbour is solid -- and leaves to the push neighbour x + c_d, slot d
// (:435-487) -- or to the cell's own slot inv(d) when the target is solid.
// FORCE_STORE adds this step's force increment to every stored value that
// lands in a cell's node view (not to ghost slots, which no node view
// covers after an odd step).  A direction whose pull neighbour lies beyond
// a non-periodic face leaves the cell's own slot d unwritten in this step
// (its writer would be the ghost cell); that slot is still part of the
// node view the reference forces after the step, so it gets the increment
// here too.
template <class Codec, bool FORCE_LOAD, bool FORCE_STORE>
__device__ __forceinline__ void aa_odd_cell(typename Codec::S * pdf, Codec cod, const AaGeom & g,
                                            const TrtC<typename Codec::R> & k, const ForceT<typename Codec::R> & ft,
                                            int x, int y, int z, uint32_t w)
{
	using R = typename Codec::R;
	const bool edge = (w & kInfoBoundary) != 0;
	const uint64_t self = aa_slot(g, 0, x, y, z);

	uint64_t src[18], dst[18];
	uint32_t ghostPull = 0;
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		const int id = inv_dir(d);
		// load of direction d
		if (w & (1u << id)) {
			src[d] = self + (uint64_t)d * g.plane;
		} else if (!edge) {
			src[d] = self + (uint64_t)id * g.plane - (int64_t)cx(d) * g.gY * g.zStride
			         - (int64_t)cy(d) * g.zStride - (int64_t)cz(d);
		} else {
			int qx = x - cx(d), qy = y - cy(d), qz = z - cz(d);
			aa_resolve(g, qx, qy, qz);
			src[d] = aa_slot(g, id, qx, qy, qz);
			if (qx < 0 || qx >= g.X || qy < 0 || qy >= g.Y || qz < 0 || qz >= g.Z) ghostPull |= 1u << d;
		}
		// store of direction d
		if (w & (1u << d)) {
			dst[d] = self + (uint64_t)id * g.plane;
		} else if (!edge) {
			dst[d] = self + (uint64_t)d * g.plane + (int64_t)cx(d) * g.gY * g.zStride
			         + (int64_t)cy(d) * g.zStride + (int64_t)cz(d);
		} else {
			int qx = x + cx(d), qy = y + cy(d), qz = z + cz(d);
			aa_resolve(g, qx, qy, qz);
			dst[d] = aa_slot(g, d, qx, qy, qz);
		}
	}
	R f[19];
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		f[d] = cod.template load<false>(&pdf[src[d]]);
		if (FORCE_LOAD) f[d] = f[d] + ft.f[d];
	}
	const uint64_t cSlot = self + (uint64_t)dC * g.plane;
	f[dC] = cod.template load<false>(&pdf[cSlot]);
	if (FORCE_LOAD) f[dC] = f[dC] + ft.f[dC];
	trt_collide(f, k);
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		R v = f[d];
		if (FORCE_STORE) {
			if (w & (1u << d)) {
				v = v + ft.f[inv_dir(d)];
			} else {
				int qx = x + cx(d), qy = y + cy(d), qz = z + cz(d);
				if (edge) aa_resolve(g, qx, qy, qz);
				const bool inside = qx >= 0 && qx < g.X && qy >= 0 && qy < g.Y && qz >= 0 && qz < g.Z;
				if (inside) v = v + ft.f[d];
			}
		}
		cod.template store<false>(&pdf[dst[d]], v);
	}
	cod.template store<false>(&pdf[cSlot], FORCE_STORE ? f[dC] + ft.f[dC] : f[dC]);
	if (FORCE_STORE && ghostPull) {
#pragma unroll
		for (int d = 0; d < 18; ++d) {
			if (ghostPull & (1u << d)) {
				const uint64_t own = self + (uint64_t)d * g.plane;
				cod.template store<false>(&pdf[own], cod.template load<false>(&pdf[own]) + ft.f[d]);
			}
		}
	}
}

template <class Codec, bool FORCE_LOAD, bool FORCE_STORE>
__global__ __launch_bounds__(512) void k_aa_odd(typename Codec::S * pdf, Codec cod, const uint32_t * __restrict__ info,
                                                AaGeom g, TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft)
{
	int x, y, z;
	if (!aa_cell(g, x, y, z)) return;
	const uint32_t w = info[((size_t)x * g.Y + y) * g.Z + z];
	if (w & kInfoFluid) aa_odd_cell<Codec, FORCE_LOAD, FORCE_STORE>(pdf, cod, g, k, ft, x, y, z, w);
}

// V stored values of consecutive z cells as one access.  The pulls and
// pushes of the odd step are offset by one cell in z, so the address is
// aligned to one stored value only: 4 bytes for the FP32 pairs, which the
// device's 8-byte global accesses accept.
template <class S, int V>
__device__ __forceinline__ VecS<S, V> ld_run(const S * p)
{
	VecS<S, V> t;
	__builtin_memcpy(&t, p, sizeof(t));
	return t;
}

template <class S, int V>
__device__ __forceinline__ void st_run(S * p, const VecS<S, V> & t)
{
	__builtin_memcpy(p, &t, sizeof(t));
}

// Odd step over groups of V consecutive z cells (FP32 pairs, -vec auto): a
// group of interior cells (fluid, no solid neighbour, not on the box
// boundary) moves each direction with one access of V values for the pull
// and one for the push; any other group goes through aa_odd_cell cell by
// cell.  Each cell's arithmetic is that of k_aa_odd.
template <class Codec, int V, bool FORCE_LOAD, bool FORCE_STORE>
__global__ __launch_bounds__(512) void k_aa_odd_v(typename Codec::S * pdf, Codec cod, const uint32_t * __restrict__ info,
                                                  AaGeom g, TrtC<typename Codec::R> k, ForceT<typename Codec::R> ft)
{
	using R = typename Codec::R;
	using S = typename Codec::S;
	const int z0 = (int)(blockIdx.x * blockDim.x + threadIdx.x) * V;
	const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
	const int x = (int)blockIdx.z;
	if (z0 >= g.Z || y >= g.Y) return;
	const size_t i0 = ((size_t)x * g.Y + y) * g.Z + z0;
	uint32_t w[V];
	bool interior = true;
#pragma unroll
	for (int j = 0; j < V; ++j) {
		w[j] = z0 + j < g.Z ? info[i0 + j] : 0u;
		interior = interior && w[j] == kInfoFluid;
	}
	if (!interior) {
#pragma unroll
		for (int j = 0; j < V; ++j)
			if (w[j] & kInfoFluid) aa_odd_cell<Codec, FORCE_LOAD, FORCE_STORE>(pdf, cod, g, k, ft, x, y, z0 + j, w[j]);
		return;
	}
	const uint64_t self = aa_slot(g, 0, x, y, z0);
	R f[V][19];
#pragma unroll
	for (int d = 0; d < 18; ++d) {
		const VecS<S, V> v = ld_run<S, V>(&pdf[self + (uint64_t)inv_dir(d) * g.plane - (int64_t)cx(d) * g.gY * g.zStride -
		                                      (int64_t)cy(d) * g.zStride - (int64_t)cz(d)]);
#pragma unroll
		for (int j = 0; j < V; ++j) f[j][d] = cod.dec(v[j]);
	}
	const uint64_t cSlot = self + (uint64_t)dC * g.plane;
	{
		const VecS<S, V> v = ld_run<S, V>(&pdf[cSlot]);
#pragma unroll
		for (int j = 0; j < V; ++j) f[j][dC] = cod.dec(v[j]);
	}
#pragma unroll
	for (int j = 0; j < V; ++j) {
		if (FORCE_LOAD) {
#pragma unroll
			for (int d = 0; d < 19; ++d) f[j][d] = f[j][d] + ft.f[d];
		}
		trt_collide(f[j], k);
	}
#pragma unroll
	for (int d = 0; d < 19; ++d) {
		// interior cells: every push target lies in the box, so the stored
		// value takes this step's increment of its own direction
		VecS<S, V> v;
#pragma unroll
		for (int j = 0; j < V; ++j) v[j] = cod.enc(FORCE_STORE ? f[j][d] + ft.f[d] : f[j][d]);
		if (d == dC) st_run<S, V>(&pdf[cSlot], v);
		else st_run<S, V>(&pdf[self + (uint64_t)d * g.plane + (int64_t)cx(d) * g.gY * g.zStride +
		                       (int64_t)cy(d) * g.zStride + (int64_t)cz(d)], v);
	}
}

// ----------------------------------------------------------- node access
// Before the first and after every Kernel() call (an even number of steps)
// the node view is local: direction d of a cell in its own slot d
// (src/BenchKernelD3Q19AaCommon.c:153-158).
static bool aa_local_layout(const AaData * A, int x, int y, int z)
{
	if (A->c.iteration % 2 == 0) return fail("aa-soa: node access after an even step is not supported by this build.");
	if (x < 0 || y < 0 || z < 0 || x >= A->g.X || y >= A->g.Y || z >= A->g.Z)
		return fail("aa-soa: node (%d, %d, %d) outside the lattice.", x, y, z);
	return true;
}

static double aa_shadow_get(const AaData * A, uint64_t i, int d)
{
	return shadow_value(A->c.args.prec, A->c.args.f16Shift, A->hPdf, (size_t)i, d);
}

static void aa_shadow_set(AaData * A, uint64_t i, int d, double v)
{
	shadow_store(A->c.args.prec, A->c.args.f16Shift, A->hPdf, (size_t)i, d, v);
}

static size_t aa_bytes(const AaData * A) { return elem_bytes(A->c.args.prec) * A->nPdf; }

// Slot that direction d of cell (x, y, z) is pulled from after an even
// step: inv(d) of the pull neighbour x - c_d (its periodic image, or the
// ghost cell on an open face), or the cell's own slot d when that neighbour
// is solid -- the source the odd step loads, and the slot the reference's
// bounce-back fix-ups fill (src/BenchKernelD3Q19Aa.c:301-311).
static uint64_t aa_pull_slot(const AaData * A, int x, int y, int z, int d)
{
	const AaGeom & g = A->g;
	if (d == dC) return aa_slot(g, d, x, y, z);
	const int id = inv_dir(d);
	if (A->hInfo[((size_t)x * g.Y + y) * g.Z + z] & (1u << id)) return aa_slot(g, d, x, y, z);
	int qx = x - cx(d), qy = y - cy(d), qz = z - cz(d);
	aa_resolve(g, qx, qy, qz);
	return aa_slot(g, id, qx, qy, qz);
}

static void aa_get_node(KernelData * kd, int x, int y, int z, PdfT * pdfs)
{
	AaData * A = as_aa(kd);
	if (A->pullView) {
		if (x < 0 || y < 0 || z < 0 || x >= A->g.X || y >= A->g.Y || z >= A->g.Z) {
			fail("aa-soa: node (%d, %d, %d) outside the lattice.", x, y, z);
			return;
		}
		for (int d = 0; d < 19; ++d) pdfs[d] = aa_shadow_get(A, aa_pull_slot(A, x, y, z, d), d) + A->viewFt[d];
		return;
	}
	if (!aa_local_layout(A, x, y, z)) return;
	for (int d = 0; d < 19; ++d) pdfs[d] = aa_shadow_get(A, aa_slot(A->g, d, x, y, z), d);
}

static void aa_set_node(KernelData * kd, int x, int y, int z, PdfT * pdfs)
{
	AaData * A = as_aa(kd);
	if (A->pullView) { fail("aa-soa: SetNode during run-time output is not supported."); return; }
	if (!aa_local_layout(A, x, y, z) || !shadow_writable(&A->c)) return;
	for (int d = 0; d < 19; ++d) aa_shadow_set(A, aa_slot(A->g, d, x, y, z), d, pdfs[d]);
	A->c.shadowDirty = 1;
}

static void aa_bc_get(KernelData * kd, int x, int y, int z, int dir, PdfT * pdf)
{
	AaData * A = as_aa(kd);
	if (!aa_local_layout(A, x, y, z)) return;
	*pdf = aa_shadow_get(A, aa_slot(A->g, dir, x, y, z), dir);
}

static void aa_bc_set(KernelData * kd, int x, int y, int z, int dir, PdfT pdf)
{
	AaData * A = as_aa(kd);
	if (!aa_local_layout(A, x, y, z) || !shadow_writable(&A->c)) return;
	aa_shadow_set(A, aa_slot(A->g, dir, x, y, z), dir, pdf);
	A->c.shadowDirty = 1;
}

// Cell sums in the local layout (after a completed Kernel() call), in FP64:
// per block of 256 cells the sum over its fluid cells of sum_d v_d and of
// sum_d c_x,d v_d (v the stored value); k_sum_pairs then adds the block sums
// in a fixed order.
template <class Codec>
__global__ __launch_bounds__(256) void k_aa_moments(const typename Codec::S * __restrict__ pdf, Codec cod,
                                                    const uint32_t * __restrict__ info, AaGeom g,
                                                    double * __restrict__ partial)
{
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	double m = 0.0, jx = 0.0;
	if (i < (uint32_t)g.X * g.Y * g.Z && (info[i] & kInfoFluid)) {
		const int z = (int)(i % g.Z), y = (int)((i / g.Z) % g.Y), x = (int)(i / ((uint32_t)g.Z * g.Y));
		const uint64_t base = aa_slot(g, 0, x, y, z);
		for (int d = 0; d < 19; ++d) {
			const double v = (double)cod.template load<false>(&pdf[base + (uint64_t)d * g.plane]);
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

// One even + odd step pair, or one of its halves.  ev: four events around
// the two launches (per-launch timing) or null.
struct AaLaunch {
	dim3 grid, block;        // odd step: one z row segment per block
	dim3 vgrid, vblock;      // odd step over pairs of z cells (FP32)
	dim3 egrid, eblock;      // even step: blockDim.y rows of y per block
	size_t elds;
	int vec;                 // cells per thread in the even step (1, or VMAX of the codec)
	bool oddPairs;           // FP32 odd step over pairs of z cells
	hipStream_t stream;
	const uint32_t * info;               // odd step
	const unsigned long long * fluid;    // even step
	AaGeom g;
	bool nt;
};

template <class Codec>
static void aa_launch_even(const AaLaunch & L, typename Codec::S * pdf, Codec cod, const TrtC<typename Codec::R> & k,
                           const ForceT<typename Codec::R> & ft, bool fl)
{
	constexpr int V = Codec::VMAX;
	if constexpr (V > 1) {
		if (L.vec == V) {
			if (fl && L.nt)  k_aa_even_v<Codec, V, true, true><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
			else if (fl)     k_aa_even_v<Codec, V, true, false><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
			else if (L.nt)   k_aa_even_v<Codec, V, false, true><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
			else             k_aa_even_v<Codec, V, false, false><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
			return;
		}
	}
	if (fl && L.nt)  k_aa_even<Codec, true, true><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
	else if (fl)     k_aa_even<Codec, true, false><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
	else if (L.nt)   k_aa_even<Codec, false, true><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
	else             k_aa_even<Codec, false, false><<<L.egrid, L.eblock, L.elds, L.stream>>>(pdf, cod, L.fluid, L.g, k, ft);
}

// doEven / doOdd: launch that half of the pair (all events of a pair are
// recorded when both halves of it are launched, in one call or two).
template <class Codec>
static void aa_launch(const AaLaunch & L, typename Codec::S * pdf, Codec cod, const TrtConsts & k0, const ForceTerms & f0,
                      bool force, bool fl, bool fs, hipEvent_t * ev, bool doEven, bool doOdd)
{
	using R = typename Codec::R;
	const TrtC<R> k = trt_cast<R>(k0);
	const ForceT<R> ft = force_cast<R>(f0);
	if (doEven) {
		if (ev) (void)hipEventRecord(ev[0], L.stream);
		aa_launch_even(L, pdf, cod, k, ft, fl);
		if (ev) (void)hipEventRecord(ev[1], L.stream);
	}
	if (!doOdd) return;
	if (ev) (void)hipEventRecord(ev[2], L.stream);
	// the odd step loads the force increment of the even step whenever the force is on
	// groups of cells only for 4-byte values: FP16 groups (8-byte accesses at
	// 2-byte alignment) measured 1.5x slower than one cell per thread
	constexpr int V = Codec::VMAX;
	if constexpr (V > 1 && sizeof(typename Codec::S) == 4) {
		if (L.oddPairs) {
			if (fs)          k_aa_odd_v<Codec, V, true, true><<<L.vgrid, L.vblock, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
			else if (force)  k_aa_odd_v<Codec, V, true, false><<<L.vgrid, L.vblock, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
			else             k_aa_odd_v<Codec, V, false, false><<<L.vgrid, L.vblock, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
			if (ev) (void)hipEventRecord(ev[3], L.stream);
			return;
		}
	}
	if (fs)          k_aa_odd<Codec, true, true><<<L.grid, L.block, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
	else if (force)  k_aa_odd<Codec, true, false><<<L.grid, L.block, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
	else             k_aa_odd<Codec, false, false><<<L.grid, L.block, 0, L.stream>>>(pdf, cod, L.info, L.g, k, ft);
	if (ev) (void)hipEventRecord(ev[3], L.stream);
}

// --------------------------------------------------------------- Kernel()
static bool aa_download(KernelData * kd)
{
	AaData * A = as_aa(kd);
	if (A->c.shadowDirty) return fail("host lattice modified through SetNode since the last Kernel() call; not overwriting it.");
	if (!hip_ok(hipMemcpyAsync(A->hPdf, A->dPdf, aa_bytes(A), hipMemcpyDeviceToHost, A->c.stream), "download of the device lattice")) return false;
	if (!hip_ok(hipStreamSynchronize(A->c.stream), "hipStreamSynchronize")) return false;
	A->c.shadowStale = 0;
	return true;
}

static bool aa_upload(AaData * A)
{
	Common & c = A->c;
	if (!c.shadowDirty) { c.t.uploadMs = 0.0; return true; }
	const double t0 = (double)clock() / CLOCKS_PER_SEC;
	if (!hip_ok(hipMemcpyAsync(A->dPdf, A->hPdf, aa_bytes(A), hipMemcpyHostToDevice, c.stream), "upload of the host lattice")) return false;
	if (!hip_ok(hipStreamSynchronize(c.stream), "hipStreamSynchronize")) return false;
	c.t.uploadMs = ((double)clock() / CLOCKS_PER_SEC - t0) * 1e3;
	c.shadowDirty = 0;
	return true;
}

// Host copy of the current state for the output functions' GetNode.
static bool aa_output_sync(AaData * A)
{
	if (A->c.shadowDirty || !A->c.shadowStale) return true;   // the host copy is current
	return aa_download(&A->kd);
}

// Threads along z of a block over rows of n cells (or cell groups), at
// most 256: the odd step splits a row into equal parts rounded up to a
// wavefront (384 cells: 2 x 192); the even step (and the FP32 odd step over
// pairs), whose blocks take several rows, takes the power of two from
// `least` to 256 that leaves the fewest idle threads in a row's last block
// (384 cells: 3 x 128; 32 FP32 pairs of a 64-cell row: 32, two rows per
// wavefront).
static int odd_row_threads(int n)
{
	const int parts = (n + 255) / 256;
	return ((n + parts - 1) / parts + 63) / 64 * 64;
}

static int even_row_threads(int n, int least)
{
	int best = 256, idle = 1 << 30;
	for (int b = 256; b >= least; b /= 2) {
		const int i = (n + b - 1) / b * b - n;
		if (i < idle) { best = b; idle = i; }
	}
	return best;
}

static bool aa_run(KernelData * kd, const CaseData * cd, LatticeDesc * ld)
{
	AaData * A = as_aa(kd);
	Common & c = A->c;
	const int steps = cd->MaxIterations > 0 ? cd->MaxIterations + (cd->MaxIterations & 1) : 0;
	if (!(cd->Omega > 0.0 && cd->Omega < 2.0)) return fail("omega must be in (0, 2), got %f.", cd->Omega);

	// Run-time output at the points of the reference's VTK_OUTPUT / STATISTICS
	// build (src/BenchKernelD3Q19Aa.c:102-112, 296-332, 513-547): before the
	// loop; after each even step, once on the state before the bounce-back
	// fix-ups (the local view of the swapped values) and once on the view at
	// even parity (pull view, force added); after each odd step.  The reference
	// tests (iter + 1 % VtkModulus) == 0 after the odd step, which holds only
	// for VtkModulus 1 at iter 0; that is kept.
	const bool vtk = ld && cd->VtkOutput;
	const bool stats = ld && c.args.statistics;
	if (vtk && !VtkWrite) return fail("VTK output requested but the program does not link VtkWrite.");
	if (stats && !KernelStatistics) return fail("-statistics requested but the program does not link KernelStatistics.");
	if (vtk && cd->VtkModulus <= 0) return fail("VtkModulus must be > 0.");
	if (stats && cd->StatisticsModulus <= 0) return fail("StatisticsModulus must be > 0.");
	CaseData * ccd = const_cast<CaseData *>(cd);
	auto due = [&](int i) { return i % 10 == 0 || i % cd->StatisticsModulus == 0; };
	if ((vtk || stats) && !A->hInfo) {
		const size_t n = (size_t)A->g.X * A->g.Y * A->g.Z;
		A->hInfo = (uint32_t *)malloc(sizeof(uint32_t) * n);
		if (!A->hInfo) return fail("allocating %zu bytes failed.", sizeof(uint32_t) * n);
		if (!hip_ok(hipMemcpy(A->hInfo, A->dInfo, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "cell info download")) {
			free(A->hInfo);
			A->hInfo = nullptr;
			return false;
		}
	}
	if (vtk || stats) {
		if (!aa_output_sync(A)) return false;
		if (vtk) VtkWrite(ld, kd, ccd, -1);
		if (stats) KernelStatistics(kd, ld, ccd, 0);
	}

	if (!aa_upload(A)) return false;

	const TrtConsts k = trt_consts(cd->Omega, c.args.magic);
	const ForceTerms ft = force_terms(cd->XForce);
	const bool force = c.args.forceOn != 0;
	const bool perLaunch = c.launchEvents != 0;
	if (perLaunch && !common_events(&c, steps)) return false;

	const AaGeom g = A->g;
	const int bs = odd_row_threads(g.Z);
	const dim3 block((unsigned)bs), grid((unsigned)((g.Z + bs - 1) / bs), (unsigned)g.Y, (unsigned)g.X);
	// V cells per thread for the narrow types (-vec auto); even step blocks
	// of 512 threads over several y rows at one block per CU, as the list
	// kernel's even step (stream_shape)
	const int vec = c.args.evenVec == 0 ? 1 : c.args.prec == PREC_FP32S ? 2 : c.args.prec == PREC_FP16S ? 4 : 1;
	const int zq = (g.Z + vec - 1) / vec;
	// groups of cells (8 bytes per lane): row segments of 16 lanes still span
	// a 128-B line; one value per lane keeps the former 64-lane minimum
	const int ebs = even_row_threads(zq, vec > 1 ? 16 : 64);
	const unsigned egx = (unsigned)((zq + ebs - 1) / ebs);
	// Defaults measured at 64^3 and 256^3 (duct channel): on lattices larger
	// than the 256 MiB last-level cache non-temporal accesses (FP64 with the
	// occupancy cap, the narrow types at full occupancy), on smaller ones plain
	// accesses with the cap (non-temporal accesses would push the
	// cache-resident lattice out).
	const bool big = aa_bytes(A) > kLastLevelCacheBytes;
	const StreamShape es = (big && c.args.prec != PREC_FP64 && !c.args.evenCapGiven)
	                       ? StreamShape{ebs, 0} : stream_shape(&c, ebs, (uint64_t)egx * ebs * g.Y * g.X);
	// at least one wavefront per block when rows are shorter than one
	const unsigned rows = (unsigned)(es.block / ebs) * (unsigned)ebs < 64u ? 64u / (unsigned)ebs
	                                                                     : (unsigned)(es.block / ebs);
	const dim3 eblock((unsigned)ebs, rows), egrid(egx, ((unsigned)g.Y + rows - 1) / rows, grid.z);
	const bool nt = c.args.nontemporal == 1 || (c.args.nontemporal == -1 && big);
	// FP32 odd step over pairs of z cells: blocks of 256 threads over 256 / ebs
	// rows of y; measured faster than one cell per thread only on lattices
	// over 256 MiB (256^3 duct 0.488 against 0.497 ms; 64^3 27.7 against
	// 15.2 us, 128^3 69.6 against 61.0 us)
	const bool oddPairs = vec == 2 && (c.args.oddPairs == 1 || (c.args.oddPairs == -1 && big));
	const unsigned vrows = (unsigned)(256 / ebs);
	const dim3 vblock((unsigned)ebs, vrows), vgrid(egx, ((unsigned)g.Y + vrows - 1) / vrows, (unsigned)g.X);
	const AaLaunch L = {grid, block, vgrid, vblock, egrid, eblock, es.lds, vec, oddPairs, c.stream, A->dInfo,
	                    A->dFluid, g, nt};
	const float f16s = (float)(1 << c.args.f16Shift);
	auto half = [&](bool on, bool fl, bool fs, int it, bool doEven, bool doOdd) {
		hipEvent_t * ev = perLaunch ? c.ev + 2 * it : nullptr;
		switch (c.args.prec) {
			case PREC_FP32S: aa_launch(L, (float *)A->dPdf, CodecF32{}, k, ft, on, fl, fs, ev, doEven, doOdd); break;
			case PREC_FP16S:
				aa_launch(L, (uint16_t *)A->dPdf, CodecF16{f16s, 1.0f / f16s}, k, ft, on, fl, fs, ev, doEven, doOdd);
				break;
			default: aa_launch(L, (double *)A->dPdf, CodecF64{}, k, ft, on, fl, fs, ev, doEven, doOdd);
		}
	};
	// the host copy after a launch: the kernels so far, then a download
	auto output_state = [&]() -> bool {
		if (!hip_ok(hipGetLastError(), "step kernel launch")) return false;
		c.shadowStale = 1;
		return aa_download(kd);
	};

	(void)hipGetLastError();   // launch checks below must see only this loop's errors
	if (!hip_ok(hipEventRecord(c.evLoop[0], c.stream), "hipEventRecord")) return false;
	bool segmentStart = true;  // the force of the previous step is in the stored values
	bool synced = true;        // the host copy holds the state of the last launch
	const bool useGraph = c.args.graph && !perLaunch && !vtk && !stats && steps > 0 && A->nFluid > 0;
	if (useGraph) {
		// the same launches as the loop below without output, replayed
		auto record = [&]() {
			for (int it = 0; it < steps; it += 2) half(force, force && it > 0, force && it + 2 >= steps, it, true, true);
		};
		if (!common_graph_run(&c, steps, cd->Omega, cd->XForce, record)) return false;
		synced = false;
	}
	for (int it = 0; it < steps && A->nFluid > 0 && !useGraph; it += 2) {
		const bool statsEven = stats && due(it), vtkEven = vtk && it % cd->VtkModulus == 0;
		const bool vtkOdd = vtk && cd->VtkModulus == 1 && it == 0, statsOdd = stats && due(it + 1);
		const bool mid = statsEven || vtkEven, end = vtkOdd || statsOdd;
		const bool last = it + 2 >= steps || end;
		synced = false;
		if (!mid) {
			half(force, force && !segmentStart, force && last, it, true, true);
		} else {
			half(force, force && !segmentStart, false, it, true, false);
			if (!output_state()) return false;
			c.iteration = it - 1;              // before the fix-ups: the swapped values in the local view
			if (statsEven) KernelStatistics(kd, ld, ccd, it);
			for (int d = 0; d < 19; ++d) A->viewFt[d] = force ? ft.f[d] : 0.0;
			A->pullView = true;
			c.iteration = it;
			if (vtkEven) VtkWrite(ld, kd, ccd, it);
			if (statsEven) KernelStatistics(kd, ld, ccd, it);
			A->pullView = false;
			half(force, false, force && last, it, false, true);
		}
		segmentStart = last;
		if (end) {
			if (!output_state()) return false;
			synced = true;
			c.iteration = it + 1;
			if (vtkOdd) VtkWrite(ld, kd, ccd, it + 1);
			if (statsOdd) KernelStatistics(kd, ld, ccd, it + 1);
		}
	}
	if (!hip_ok(hipGetLastError(), "step kernel launch")) return false;
	if (!hip_ok(hipEventRecord(c.evLoop[1], c.stream), "hipEventRecord")) return false;
	if (!hip_ok(hipStreamSynchronize(c.stream), "step kernels")) return false;

	common_collect(&c, A->nFluid > 0 ? steps : 0, perLaunch);
	c.t.steps = steps;
	kd->Duration = c.t.loopMs / 1e3;
	if (steps > 0) c.iteration = steps - 1;

	if (!synced) c.shadowStale = A->nFluid > 0 && steps > 0;
	if (c.shadowDownload || vtk) {
		const double t0 = (double)clock() / CLOCKS_PER_SEC;
		if (!aa_output_sync(A)) return false;
		c.t.downloadMs = ((double)clock() / CLOCKS_PER_SEC - t0) * 1e3;
	}
	if (vtk) VtkWrite(ld, kd, ccd, cd->MaxIterations);
	return true;
}

static void aa_kernel(LatticeDesc * ld, KernelData * kd, CaseData * cd)
{
	(void)aa_run(kd, cd, ld);
}


Common * aa_common(KernelData * kd)
{
	return (kd && kd->Kernel == aa_kernel) ? &as_aa(kd)->c : nullptr;
}

bool aa_steps(KernelData * kd, const CaseData * cd)
{
	if (!aa_common(kd)) return fail("not an aa-soa kernel.");
	return aa_run(kd, cd, nullptr);
}

bool aa_download_shadow(KernelData * kd)
{
	if (!aa_common(kd)) return fail("not an aa-soa kernel.");
	return aa_download(kd);
}

// Mass and x momentum of the device lattice in FP64 (shifted storage: the
// stored deviations plus the weights, which add 1 per fluid cell to the
// mass and nothing to the momentum).
bool aa_moments(KernelData * kd, double * mass, double * jx)
{
	if (!aa_common(kd)) return fail("not an aa-soa kernel.");
	AaData * A = as_aa(kd);
	Common & c = A->c;
	*mass = 0.0;
	*jx = 0.0;
	if (!aa_upload(A)) return false;
	if (A->nFluid == 0) return true;
	const AaGeom g = A->g;
	const uint32_t nb = (uint32_t)(((uint64_t)g.X * g.Y * g.Z + 255) / 256);
	if (!A->dMoments) {
		if (!hip_ok(hipMalloc((void **)&A->dMoments, sizeof(double) * 2 * ((size_t)nb + 1)), "allocating moment sums")) return false;
		A->momentBlocks = nb;
	}
	double * part = A->dMoments, * sum = A->dMoments + 2 * (size_t)nb;
	const float f16s = (float)(1 << c.args.f16Shift);
	(void)hipGetLastError();
	switch (c.args.prec) {
		case PREC_FP32S: k_aa_moments<<<nb, 256, 0, c.stream>>>((const float *)A->dPdf, CodecF32{}, A->dInfo, g, part); break;
		case PREC_FP16S:
			k_aa_moments<<<nb, 256, 0, c.stream>>>((const uint16_t *)A->dPdf, CodecF16{f16s, 1.0f / f16s}, A->dInfo, g, part);
			break;
		default: k_aa_moments<<<nb, 256, 0, c.stream>>>((const double *)A->dPdf, CodecF64{}, A->dInfo, g, part);
	}
	k_sum_pairs<256><<<1, 256, 0, c.stream>>>(part, nb, sum);
	if (!hip_ok(hipGetLastError(), "moment kernels")) return false;
	double h[2];
	if (!hip_ok(hipMemcpyAsync(h, sum, sizeof(h), hipMemcpyDeviceToHost, c.stream), "moment download")) return false;
	if (!hip_ok(hipStreamSynchronize(c.stream), "moment kernels")) return false;
	*mass = h[0] + (c.args.prec == PREC_FP64 ? 0.0 : (double)A->nFluid);
	*jx = h[1];
	return true;
}

static void aa_release(AaData * A)
{
	if (!A) return;
	(void)hipFree(A->dMoments);
	(void)hipFree(A->dInfo);
	(void)hipFree(A->dFluid);
	(void)hipFree(A->dPdf);
	host_free(A->hPdf);
	free(A->hInfo);
	common_free(&A->c);
	delete A;
}

}  // namespace lbm

using namespace lbm;

extern "C" {

void D3Q19AaInit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params)
{
	*kernelData = nullptr;
	AaData * A = new (std::nothrow) AaData();
	if (!A) { fail("allocating kernel data failed."); return; }
	memset(&A->kd, 0, sizeof(A->kd));
	if (!common_init(&A->c, params, false)) { aa_release(A); return; }

	AaGeom & g = A->g;
	g.X = ld->Dims[0]; g.Y = ld->Dims[1]; g.Z = ld->Dims[2];
	g.px = ld->PeriodicX; g.py = ld->PeriodicY; g.pz = ld->PeriodicZ;
	if (g.X <= 0 || g.Y <= 0 || g.Z <= 0) { fail("empty dimensions."); aa_release(A); return; }
	// the step launches take y and x as grid dimensions
	if (g.X > 65535 || g.Y > 65535) {
		fail("aa-soa: x and y extents above 65535 are not supported (launch grid limit), got %d x %d.", g.X, g.Y);
		aa_release(A);
		return;
	}
	const uint64_t nTotal = (uint64_t)g.X * g.Y * g.Z;
	if (nTotal > 0xffffffffull / 19) { fail("number of PDFs exceed 2^32."); aa_release(A); return; }
	if (!ld->Lattice) { fail("LatticeDesc without a mask."); aa_release(A); return; }
	g.gY = g.Y + 2;
	{
		const int line = (int)(128 / elem_bytes(A->c.args.prec));   // stored values per 128 B
		bool zFaceFluid = false;
		if (!g.pz) {
			const size_t face = (size_t)g.X * g.Y;
			for (size_t i = 0; i < face && !zFaceFluid; ++i)
				zFaceFluid = ld->Lattice[i] != LBM_CELL_OBSTACLE ||
				             ld->Lattice[(size_t)(g.Z - 1) * face + i] != LBM_CELL_OBSTACLE;
		}
		// compact rows: no fluid cell reaches a z ghost cell.  Measured at 256^3
		// (duct channel): even step FP64 0.94 -> 0.83 ms, FP32 0.47 -> 0.45 ms,
		// but FP16 0.22 -> 0.23 ms, so FP16 keeps the ghost cells.
		if (!zFaceFluid && elem_bytes(A->c.args.prec) >= 4) {
			g.zOff = 0;
			g.zStride = (g.Z + line - 1) / line * line;
		} else {
			g.zOff = line;
			g.zStride = ((line + g.Z + 1 + line - 1) / line) * line;
		}
	}
	g.zWords = (g.Z + 63) / 64;
	g.plane = (uint64_t)(g.X + 2) * g.gY * g.zStride;
	A->nPdf = (size_t)19 * g.plane;
	A->nFluid = ld->nFluid;

	if (!A->c.args.quiet) {
		printf("# blocking x: %3d y: %3d z: %3d\n", g.X + 2, g.Y + 2, g.Z + 2);
		printf("# allocating data for %llu LB nodes with padding (%llu bytes = %f MiB for the single lattice)\n",
		       (unsigned long long)(g.plane), (unsigned long long)aa_bytes(A), aa_bytes(A) / 1024.0 / 1024.0);
		printf("# z rows: %s, %d values apart, z = 0 at value %d\n", g.zOff ? "with ghost cells" : "compact", g.zStride,
		       g.zOff);
		fflush(stdout);
	}

	int * dMask = nullptr;
	uint32_t * dCount = nullptr;
	uint32_t hCount = 0;
	bool ok = false;
	hipStream_t s = A->c.stream;
	do {
		if (!hip_ok(hipMalloc((void **)&dMask, sizeof(int) * nTotal), "allocating geometry mask")) break;
		if (!hip_ok(hipMalloc((void **)&dCount, sizeof(uint32_t)), "allocating counter")) break;
		if (!hip_ok(hipMalloc((void **)&A->dInfo, sizeof(uint32_t) * nTotal), "allocating cell info")) break;
		const uint64_t nWords = (uint64_t)g.X * g.Y * g.zWords;
		if (!hip_ok(hipMalloc((void **)&A->dFluid, sizeof(unsigned long long) * nWords), "allocating fluid mask")) break;
		if (!hip_ok(hipMalloc((void **)&A->dPdf, aa_bytes(A)), "allocating device lattice")) break;
		if (!hip_ok(hipMemcpyAsync(dMask, ld->Lattice, sizeof(int) * nTotal, hipMemcpyHostToDevice, s), "mask upload")) break;
		if (!hip_ok(hipMemsetAsync(dCount, 0, sizeof(uint32_t), s), "hipMemset")) break;
		(void)hipGetLastError();
		k_aa_info<<<(unsigned)((nTotal + 255) / 256), 256, 0, s>>>(dMask, g, A->dInfo, dCount);
		k_aa_fluid_bits<<<(unsigned)((nWords + 255) / 256), 256, 0, s>>>(A->dInfo, g, A->dFluid);
		if (!hip_ok(hipGetLastError(), "cell info kernel")) break;
		if (!hip_ok(hipMemcpyAsync(&hCount, dCount, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "count download")) break;
		A->hPdf = host_alloc(aa_bytes(A));
		if (!A->hPdf) { fail("allocating host lattice failed: %zu bytes.", aa_bytes(A)); break; }
		// every slot starts as f = 0, ghost cells included; shifted storage
		// holds that as the code of -w_d
		if (A->c.args.prec == PREC_FP64) {
			memset(A->hPdf, 0, aa_bytes(A));
		} else {
			for (int d = 0; d < 19; ++d)
				for (uint64_t i = 0; i < g.plane; ++i) aa_shadow_set(A, (uint64_t)d * g.plane + i, d, 0.0);
		}
		if (!hip_ok(hipMemcpyAsync(A->dPdf, A->hPdf, aa_bytes(A), hipMemcpyHostToDevice, s), "lattice upload")) break;
		if (!hip_ok(hipStreamSynchronize(s), "init")) break;
		ok = true;
	} while (0);
	(void)hipFree(dMask);
	(void)hipFree(dCount);
	if (!ok) { aa_release(A); return; }
	A->nBounce = (int)hCount;
	if (!A->c.args.quiet) printf("# bounce back pdfs: %d (resolved in the odd step, no index lists)\n", A->nBounce);

	KernelData * kd = &A->kd;
	kd->Pdfs[0] = (PdfT *)A->hPdf;     // storage type given by -prec (double by default)
	kd->PdfsActive = (PdfT *)A->hPdf;
	for (int i = 0; i < 3; ++i) { kd->Dims[i] = ld->Dims[i]; kd->GlobalDims[i] = ld->Dims[i] + 2; kd->Offsets[i] = 1; }
	kd->nBounceBackPdfs = A->nBounce;
	kd->LoopBalance = 2.0 * 19 * (double)elem_bytes(A->c.args.prec) + 4.0;
	kd->GetNode = aa_get_node;
	kd->SetNode = aa_set_node;
	kd->BoundaryConditionsGetPdf = aa_bc_get;
	kd->BoundaryConditionsSetPdf = aa_bc_set;
	kd->Kernel = aa_kernel;
	A->c.download = aa_download;
	*kernelData = kd;
}

void D3Q19AaDeinit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData)
{
	(void)ld;
	if (kernelData && *kernelData) {
		aa_release(as_aa(*kernelData));
		*kernelData = nullptr;
	}
}

}  // extern "C"
