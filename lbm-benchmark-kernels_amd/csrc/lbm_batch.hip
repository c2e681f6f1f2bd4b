// lbm_batch.hip -- many list-aa-pv-soa lattices stepped together on one GPU.
//
// No counterpart in the reference, which runs one lattice per process.  Each
// lattice ("domain") is an ordinary list-aa-pv-soa kernel instance (its own
// list build, adjacency, device PDFs and host shadow, so GetNode/SetNode and
// every other entry point work on it); the batch adds:
//
//   * one step kernel per half-step for all active domains: a persistent
//     grid walks a device-resident list of chunks (domain, first node), each
//     chunk one block's worth of consecutive nodes of one domain, so small
//     domains fill the GPU together and the per-domain constants (omega,
//     force, pointers, sizes) come from a device table;
//   * the step loop of one check interval captured as HIP graphs -- one for
//     the whole chunk list, or one per cache-sized group of domains -- and
//     replayed; the chunk lists and each launch's chunk range live in device
//     memory, so domains that finish or join leave the graphs valid (the host
//     rewrites the lists and ranges between replays);
//   * FP64 per-domain moments after every interval (segmented reduction in a
//     fixed order) and a per-domain convergence test on the x momentum;
//   * an autotuner over block sizes, even-step occupancy and nodes per lane.
//
// The per-node arithmetic is lbm_list.h's even_node / even_nodes / odd_node,
// the same code the single-lattice kernels run, so a domain stepped in a
// batch ends bit-identical to the same lattice stepped alone.
#include "lbm_list.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

namespace lbm {

template <class R>
struct BatchDom {
	void * pdf;
	const unsigned long long * runMask;
	const uint32_t * runBase;
	const uint32_t * runAdj;
	const uint32_t * adj;
	uint32_t nFluid, nCells, runStride, adjStride;
	TrtC<R> k;
	ForceT<R> ft;
};

struct Chunk {
	uint32_t dom;     // index into the domain table
	uint32_t node0;   // first node (a multiple of the chunk size, hence of 64)
};

// ------------------------------------------------------------- kernels
// Persistent grids over the chunk range {first, count} read from `range`:
// block b takes chunks first + b, first + b + gridDim.x, ... of the range.
template <class Codec, int V, bool FORCE_LOAD, bool NT>
__global__ __launch_bounds__(512) void k_batch_even(const BatchDom<typename Codec::R> * __restrict__ doms,
                                                    const Chunk * __restrict__ chunks,
                                                    const uint32_t * __restrict__ range, Codec cod)
{
	using R = typename Codec::R;
	using S = typename Codec::S;
	const Chunk * __restrict__ list = chunks + range[0];
	const uint32_t nc = range[1];
	for (uint32_t c = blockIdx.x; c < nc; c += gridDim.x) {
		const Chunk ch = list[c];
		const BatchDom<R> & D = doms[ch.dom];
		const uint32_t n0 = ch.node0 + threadIdx.x * V;
		if (n0 >= D.nFluid) continue;
		if constexpr (V == 1) even_node<Codec, FORCE_LOAD, NT>((S *)D.pdf, cod, n0, D.nCells, D.k, D.ft);
		else even_nodes<Codec, V, FORCE_LOAD, NT>((S *)D.pdf, cod, n0, D.nFluid, D.nCells, D.k, D.ft);
	}
}

template <class Codec, bool FORCE_LOAD, bool FORCE_STORE, bool RLE>
__global__ __launch_bounds__(512) void k_batch_odd(const BatchDom<typename Codec::R> * __restrict__ doms,
                                                   const Chunk * __restrict__ chunks,
                                                   const uint32_t * __restrict__ range, Codec cod, uint32_t gShift)
{
	using R = typename Codec::R;
	using S = typename Codec::S;
	// the grid is a multiple of 8 << gShift blocks: each round of gridDim.x
	// chunks is handed out in the XCD-grouped order, 2^gShift consecutive
	// chunks per XCD turn
	const Chunk * __restrict__ list = chunks + range[0];
	const uint32_t nc = range[1], slot = xcd_grouped_block(blockIdx.x, gShift);
	for (uint32_t base = 0; base < nc; base += gridDim.x) {
		const uint32_t c = base + slot;
		if (c >= nc) break;
		const Chunk ch = list[c];
		const BatchDom<R> & D = doms[ch.dom];
		const uint32_t n = ch.node0 + threadIdx.x;
		if (n >= D.nFluid) continue;
		uint32_t a[18];
		if constexpr (RLE) AdjRuns{D.runMask, D.runBase, D.runAdj, D.runStride}.fetch(n, a);
		else AdjFull{D.adj, D.adjStride}.fetch(n, a);
		odd_node<Codec, FORCE_LOAD, FORCE_STORE>((S *)D.pdf, cod, a, n, D.nCells, D.k, D.ft);
	}
}

// Sum over a chunk of 256 nodes of sum_d v_d and sum_d c_x,d v_d (v the
// stored value), in FP64.
template <class Codec>
__global__ __launch_bounds__(256) void k_batch_moments(const BatchDom<typename Codec::R> * __restrict__ doms,
                                                       const Chunk * __restrict__ chunks, Codec cod,
                                                       double * __restrict__ partial)
{
	using R = typename Codec::R;
	using S = typename Codec::S;
	const Chunk ch = chunks[blockIdx.x];
	const BatchDom<R> & D = doms[ch.dom];
	const uint32_t n = ch.node0 + threadIdx.x;
	double m = 0.0, jx = 0.0;
	if (n < D.nFluid) {
		const S * pdf = (const S *)D.pdf;
		for (int d = 0; d < 19; ++d) {
			const double v = (double)cod.template load<false>(&pdf[(size_t)d * D.nCells + n]);
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

// Block i adds the chunk sums of segment [seg[i], seg[i+1]) in a fixed order.
__global__ __launch_bounds__(256) void k_batch_reduce(const uint32_t * __restrict__ seg, const double * __restrict__ partial,
                                                      double * __restrict__ out)
{
	const uint32_t b = seg[blockIdx.x], e = seg[blockIdx.x + 1];
	double m = 0.0, jx = 0.0;
	for (uint32_t i = b + threadIdx.x; i < e; i += 256) {
		m += partial[2 * i];
		jx += partial[2 * i + 1];
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
		out[2 * blockIdx.x] = sm[0];
		out[2 * blockIdx.x + 1] = sj[0];
	}
}

// f = w_d in every fluid slot of an FP64 lattice (the local layout).
__global__ __launch_bounds__(256) void k_fill_weights(double * __restrict__ pdf, uint32_t nFluid, uint32_t nCells)
{
	const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= 19ull * nFluid) return;
	const uint32_t d = (uint32_t)(i / nFluid), n = (uint32_t)(i % nFluid);
	const double w = d == dC ? 1.0 / 3.0 : (cx(d) * cx(d) + cy(d) * cy(d) + cz(d) * cz(d) == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
	pdf[(size_t)d * nCells + n] = w;
}

// ---------------------------------------------------------------- host
struct BatchEntry {
	KernelData * kd = nullptr;
	ListData * L = nullptr;
	int state = 0;              // 0 free slot, 1 active, 2 finished
	double omega = 1.0, xForce = 0.0;
	int maxSteps = 0;
	LbmHipBatchStatus st = {};
	bool havePrev = false;
	double jxPrev = 0.0;
};

// Default group size: the device's last-level cache (256 MiB on MI355X);
// 256 MiB measured best for FP64 and within 1 % of the best for shifted FP32
// on the 512-sample porous batch.
constexpr int kDefaultGroupMB = 256;

struct Batch {
	std::vector<std::string> words;
	std::vector<char *> argv;
	Parameters params = {};
	KernelArgs args;
	int device = 0, nCU = 0, ldsPerCU = 0, ldsPerBlock = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev[2] = {nullptr, nullptr};
	std::vector<BatchEntry> e;
	LbmHipBatchTuning tune = {};

	// device tables (sized on rebuild; reallocation invalidates the graphs)
	void * dDoms = nullptr;
	size_t domsBytes = 0;
	Chunk * dEven = nullptr, * dOdd = nullptr, * dMom = nullptr;
	size_t evenCapN = 0, oddCapN = 0, momCapN = 0;
	uint32_t * dCount = nullptr;  // {0, even chunks, 0, odd chunks}: the ranges of all chunks
	uint32_t * dSeg = nullptr;
	size_t segCapN = 0;
	double * dPartial = nullptr, * dSums = nullptr;
	size_t sumCapN = 0;
	std::vector<Chunk> hEven, hOdd, hMom;
	std::vector<uint32_t> hSeg;
	std::vector<int> activeIds;    // domain id of moment segment i
	std::vector<double> hSums;
	bool tablesDirty = true;
	uint32_t evenGrid = 1, oddGrid = 1;   // persistent grids over all chunks
	bool allNt = true;             // even-step access kind over all chunks
	size_t evenLds = 0;

	// tune.groupMB > 0: the active domains in groups of about that many bytes
	// (ranges of the chunk lists), each stepped through a whole interval
	// before the next, so that a group's lattices stay in the last-level cache
	struct Group { uint32_t e0, e1, o0, o1, eGrid, oGrid; bool nt; };
	std::vector<Group> groups;
	uint32_t * dGroupRange = nullptr;  // [4g .. 4g+3]: first even chunk, even chunks, first odd chunk, odd chunks
	size_t groupCapN = 0;

	// One captured graph per launch range (slot 0 the whole chunk list, or
	// slot g group g): `steps` steps over that range.  A slot is replayed
	// while its range pointer, step count, grids and even-step access kind
	// still hold; the range contents change freely between replays.
	struct Slot {
		hipGraph_t graph = nullptr;
		hipGraphExec_t exec = nullptr;
		const uint32_t * range = nullptr;
		uint32_t eGrid = 0, oGrid = 0;
		int steps = 0;
		bool nt = false;
	};
	std::vector<Slot> slots;
	bool graphDirty = true;        // drop every slot's graph before the next interval
	size_t lastSlots = 0;
	long long captures = 0, replays = 0;
	int tableVec = 1;              // nodes per thread the even chunk list was cut for

	double gpuMs = 0.0, flups = 0.0;
	int pendingSteps = 0;          // an interval launched and not yet waited for
	double * hSumsPinned = nullptr;
	size_t hSumsCap = 0;
};

static Batch * as_batch(LbmHipBatch * b) { return reinterpret_cast<Batch *>(b); }

// log2 of the odd step's chunks per XCD turn (LbmHipBatchTuning.oddGroup, a
// power of two; 0 means kXcdGroup).
static uint32_t odd_group_shift(const Batch * b)
{
	const uint32_t g = b->tune.oddGroup > 0 ? (uint32_t)b->tune.oddGroup : kXcdGroup;
	uint32_t s = 0;
	while ((1u << s) < g) ++s;
	return s;
}

static float f16_scale(const Batch * b) { return (float)(1 << b->args.f16Shift); }

static int vec_max(int prec) { return prec == PREC_FP32S ? 2 : prec == PREC_FP16S ? 4 : 1; }

// Nodes per thread of the batch's even step: the tuned value when every
// active domain's plane stride allows aligned vectors of it, else 1.
static int batch_vec(const Batch * b)
{
	int v = b->tune.vec > 0 ? b->tune.vec : 1;
	if (v > vec_max(b->args.prec)) v = 1;
	for (const BatchEntry & x : b->e)
		if (x.state == 1 && x.L->nCells % v) return 1;
	return v;
}

static bool grow(void ** p, size_t * capN, size_t needN, size_t elem, const char * what, bool * moved)
{
	if (needN <= *capN && *p) return true;
	(void)hipFree(*p);
	*p = nullptr;
	const size_t n = needN + needN / 4 + 64;
	const hipError_t err = hipMalloc(p, n * elem);
	if (err != hipSuccess) { *capN = 0; return fail("allocating %s failed: %zu bytes (%s).", what, n * elem, hipGetErrorString(err)); }
	*capN = n;
	*moved = true;
	return true;
}

template <class R>
static void fill_doms(Batch * b, std::vector<unsigned char> & out)
{
	out.assign(sizeof(BatchDom<R>) * b->e.size(), 0);
	BatchDom<R> * t = reinterpret_cast<BatchDom<R> *>(out.data());
	for (size_t i = 0; i < b->e.size(); ++i) {
		const BatchEntry & x = b->e[i];
		if (!x.L) continue;
		const ListData * L = x.L;
		BatchDom<R> & D = t[i];
		D.pdf = L->dPdf;
		D.runMask = L->dRunMask;
		D.runBase = L->dRunBase;
		D.runAdj = L->dRunAdj;
		D.adj = L->dAdj;
		D.nFluid = (uint32_t)L->nFluid;
		D.nCells = (uint32_t)L->nCells;
		D.runStride = (uint32_t)L->runStride;
		D.adjStride = (uint32_t)L->adjStride;
		D.k = trt_cast<R>(trt_consts(x.omega, b->args.magic));
		D.ft = force_cast<R>(force_terms(x.xForce));
	}
}

// Step kernel of V nodes per thread, or of one when V exceeds the codec's.
template <class Codec, int V>
static const void * even_fn()
{
	if constexpr (V <= Codec::VMAX) return (const void *)k_batch_even<Codec, V, true, true>;
	else return (const void *)k_batch_even<Codec, 1, true, true>;
}

template <class Codec>
static bool even_occupancy(int v, int block, size_t lds, int * perCU)
{
	const void * fn = v == 4 ? even_fn<Codec, 4>() : v == 2 ? even_fn<Codec, 2>() : even_fn<Codec, 1>();
	return hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(perCU, fn, block, lds), "occupancy query");
}

// Chunk lists of the active domains, the domain table and the moment
// segments; grid sizes of the persistent launches.
static bool rebuild_tables(Batch * b)
{
	const int v = batch_vec(b);
	const uint32_t eSize = (uint32_t)(b->tune.evenBlock * v), oSize = (uint32_t)b->tune.oddBlock;
	// the captured even launches must take as many nodes per thread as the
	// chunks are cut for (the set of active domains decides it)
	if (v != b->tableVec) b->graphDirty = true;
	b->tableVec = v;
	b->hEven.clear();
	b->hOdd.clear();
	b->hMom.clear();
	b->hSeg.assign(1, 0);
	b->activeIds.clear();
	for (size_t i = 0; i < b->e.size(); ++i) {
		const BatchEntry & x = b->e[i];
		if (x.state != 1) continue;
		const uint32_t nF = (uint32_t)x.L->nFluid;
		for (uint32_t n = 0; n < nF; n += eSize) b->hEven.push_back({(uint32_t)i, n});
		for (uint32_t n = 0; n < nF; n += oSize) b->hOdd.push_back({(uint32_t)i, n});
		for (uint32_t n = 0; n < nF; n += 256) b->hMom.push_back({(uint32_t)i, n});
		b->hSeg.push_back((uint32_t)b->hMom.size());
		b->activeIds.push_back((int)i);
	}
	// Even-step accesses: forced by tune.nontemporal (0 / 1), or (-1) per
	// launch range non-temporal only when what a step touches there -- the
	// fluid nodes' PDFs and the odd step's adjacency -- exceeds the
	// last-level cache (a range that fits stays cached from step to step and
	// streams faster with plain accesses)
	auto touched = [&](const ListData * L) {
		return (double)elem_bytes(b->args.prec) * 19 * L->nFluid +
		       (b->args.adjRle ? 72.0 * L->nRunsWave + 12.0 * L->nWaves : 72.0 * L->nFluid);
	};
	auto range_nt = [&](double t) {
		return b->tune.nontemporal < 0 ? t > (double)kLastLevelCacheBytes : b->tune.nontemporal != 0;
	};
	double allTouched = 0.0;
	for (const BatchEntry & x : b->e)
		if (x.state == 1) allTouched += touched(x.L);
	b->allNt = range_nt(allTouched);
	b->groups.clear();
	if (b->tune.groupMB > 0) {
		const double budget = (double)b->tune.groupMB * 1048576.0;
		Batch::Group cur = {0, 0, 0, 0, 0, 0, false};
		double bytes = 0.0, inGroup = 0.0;
		for (size_t i = 0; i < b->e.size(); ++i) {
			const BatchEntry & x = b->e[i];
			if (x.state != 1) continue;
			const ListData * L = x.L;
			const double db = (double)elem_bytes(b->args.prec) * 19 * L->nCells +
			                  (b->args.adjRle ? 72.0 * L->runStride + 12.0 * L->nWaves : 72.0 * L->adjStride);
			if (cur.e1 > cur.e0 && bytes + db > budget) {
				cur.nt = range_nt(inGroup);
				b->groups.push_back(cur);
				cur = {cur.e1, cur.e1, cur.o1, cur.o1, 0, 0, false};
				bytes = inGroup = 0.0;
			}
			const uint32_t nF = (uint32_t)L->nFluid;
			cur.e1 += (nF + eSize - 1) / eSize;
			cur.o1 += (nF + oSize - 1) / oSize;
			bytes += db;
			inGroup += touched(L);
		}
		if (cur.e1 > cur.e0) {
			cur.nt = range_nt(inGroup);
			b->groups.push_back(cur);
		}
		// one group is the whole chunk list
		if (b->groups.size() == 1) b->groups.clear();
	}
	bool moved = false;
	std::vector<unsigned char> doms;
	if (b->args.prec == PREC_FP64) fill_doms<double>(b, doms);
	else fill_doms<float>(b, doms);
	size_t domsCapN = b->domsBytes;
	if (!grow(&b->dDoms, &domsCapN, doms.size(), 1, "domain table", &moved)) return false;
	b->domsBytes = domsCapN;
	if (!grow((void **)&b->dEven, &b->evenCapN, b->hEven.size(), sizeof(Chunk), "even chunk list", &moved)) return false;
	if (!grow((void **)&b->dOdd, &b->oddCapN, b->hOdd.size(), sizeof(Chunk), "odd chunk list", &moved)) return false;
	if (!grow((void **)&b->dMom, &b->momCapN, b->hMom.size(), sizeof(Chunk), "moment chunk list", &moved)) return false;
	if (!grow((void **)&b->dSeg, &b->segCapN, b->hSeg.size(), sizeof(uint32_t), "moment segments", &moved)) return false;
	const size_t sumsN = 2 * (b->hMom.size() > b->activeIds.size() ? b->hMom.size() : b->activeIds.size());
	size_t cap = b->sumCapN;
	if (!grow((void **)&b->dPartial, &cap, sumsN, sizeof(double), "moment sums", &moved)) return false;
	cap = b->sumCapN;
	if (!grow((void **)&b->dSums, &cap, sumsN, sizeof(double), "moment sums", &moved)) return false;
	b->sumCapN = cap;
	if (!b->dCount && !hip_ok(hipMalloc((void **)&b->dCount, 4 * sizeof(uint32_t)), "allocating chunk counts")) return false;
	if (!b->groups.empty() &&
	    !grow((void **)&b->dGroupRange, &b->groupCapN, 4 * b->groups.size(), sizeof(uint32_t), "group chunk ranges", &moved))
		return false;
	if (moved) b->graphDirty = true;

	const uint32_t counts[4] = {0, (uint32_t)b->hEven.size(), 0, (uint32_t)b->hOdd.size()};
	hipStream_t s = b->stream;
	if (!doms.empty() && !hip_ok(hipMemcpyAsync(b->dDoms, doms.data(), doms.size(), hipMemcpyHostToDevice, s), "table upload")) return false;
	if (!b->hEven.empty() && !hip_ok(hipMemcpyAsync(b->dEven, b->hEven.data(), sizeof(Chunk) * b->hEven.size(), hipMemcpyHostToDevice, s), "chunk upload")) return false;
	if (!b->hOdd.empty() && !hip_ok(hipMemcpyAsync(b->dOdd, b->hOdd.data(), sizeof(Chunk) * b->hOdd.size(), hipMemcpyHostToDevice, s), "chunk upload")) return false;
	if (!b->hMom.empty() && !hip_ok(hipMemcpyAsync(b->dMom, b->hMom.data(), sizeof(Chunk) * b->hMom.size(), hipMemcpyHostToDevice, s), "chunk upload")) return false;
	if (!hip_ok(hipMemcpyAsync(b->dSeg, b->hSeg.data(), sizeof(uint32_t) * b->hSeg.size(), hipMemcpyHostToDevice, s), "segment upload")) return false;
	if (!hip_ok(hipMemcpyAsync(b->dCount, counts, sizeof(counts), hipMemcpyHostToDevice, s), "count upload")) return false;
	std::vector<uint32_t> gRanges;
	for (const Batch::Group & g : b->groups) {
		gRanges.push_back(g.e0);
		gRanges.push_back(g.e1 - g.e0);
		gRanges.push_back(g.o0);
		gRanges.push_back(g.o1 - g.o0);
	}
	if (!gRanges.empty() &&
	    !hip_ok(hipMemcpyAsync(b->dGroupRange, gRanges.data(), sizeof(uint32_t) * gRanges.size(), hipMemcpyHostToDevice, s),
	            "range upload"))
		return false;
	if (!hip_ok(hipStreamSynchronize(s), "table upload")) return false;

	// persistent grids: the blocks that fit on the device at once
	// the single-lattice even step's occupancy cap (stream_shape)
	const size_t want = (size_t)b->ldsPerCU / 2 + 8192;
	b->evenLds = (b->tune.evenCap && b->ldsPerCU > 0 && want <= (size_t)b->ldsPerBlock) ? want : 0;
	int perCU = 1;
	const int blk = b->tune.evenBlock;
	bool ok;
	switch (b->args.prec) {
		case PREC_FP32S: ok = even_occupancy<CodecF32>(v, blk, b->evenLds, &perCU); break;
		case PREC_FP16S: ok = even_occupancy<CodecF16>(v, blk, b->evenLds, &perCU); break;
		default: ok = even_occupancy<CodecF64>(1, blk, b->evenLds, &perCU);
	}
	if (!ok) return false;
	if (perCU < 1) perCU = 1;
	const uint32_t eg = (uint32_t)(perCU * b->nCU), og0 = (uint32_t)((2048 / b->tune.oddBlock) * b->nCU);
	int oddPerCU = 1;
	const void * oddFn = b->args.prec == PREC_FP64 ? (const void *)k_batch_odd<CodecF64, true, true, true>
	                   : b->args.prec == PREC_FP32S ? (const void *)k_batch_odd<CodecF32, true, true, true>
	                   : (const void *)k_batch_odd<CodecF16, true, true, true>;
	if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&oddPerCU, oddFn, b->tune.oddBlock, 0), "occupancy query")) return false;
	const uint32_t og = oddPerCU > 0 ? (uint32_t)(oddPerCU * b->nCU) : og0;
	const uint32_t ne = counts[1] ? counts[1] : 1, no = counts[3] ? counts[3] : 1;
	const uint32_t groupBlocks = 8u << odd_group_shift(b);
	b->evenGrid = eg < ne ? eg : ne;
	b->oddGrid = ((og < no ? og : no) + groupBlocks - 1) / groupBlocks * groupBlocks;
	for (Batch::Group & g : b->groups) {
		const uint32_t gne = g.e1 - g.e0, gno = g.o1 - g.o0;
		g.eGrid = eg < gne ? eg : gne;
		g.oGrid = ((og < gno ? og : gno) + groupBlocks - 1) / groupBlocks * groupBlocks;
	}
	b->tablesDirty = false;
	return true;
}

// The chunks one launch walks: the even and odd chunk lists, the device
// words {first, count} of the range taken from each, and the persistent grid
// sizes the range needs now.
struct ChunkRange {
	const Chunk * even;
	const Chunk * odd;
	const uint32_t * evenRange;
	const uint32_t * oddRange;
	uint32_t evenGrid, oddGrid;
	bool nt;                    // non-temporal even-step accesses
};

static ChunkRange all_chunks(const Batch * b)
{
	return {b->dEven, b->dOdd, b->dCount, b->dCount + 2, b->evenGrid, b->oddGrid, b->allNt};
}

static ChunkRange group_chunks(const Batch * b, size_t g)
{
	const Batch::Group & G = b->groups[g];
	return {b->dEven, b->dOdd, b->dGroupRange + 4 * g, b->dGroupRange + 4 * g + 2, G.eGrid, G.oGrid, G.nt};
}

// Launch range of slot g: the whole chunk list, or group g.
static size_t slot_count(const Batch * b) { return b->groups.empty() ? 1 : b->groups.size(); }
static ChunkRange slot_chunks(const Batch * b, size_t g) { return b->groups.empty() ? all_chunks(b) : group_chunks(b, g); }

template <class Codec, int V>
static void launch_batch_even(Batch * b, const ChunkRange & c, Codec cod, bool fl)
{
	static_assert(V <= Codec::VMAX, "more nodes per access than the storage type packs");
	using R = typename Codec::R;
	const BatchDom<R> * doms = (const BatchDom<R> *)b->dDoms;
	const dim3 g(c.evenGrid), t(b->tune.evenBlock);
	const bool nt = c.nt;
	if (fl && nt)  k_batch_even<Codec, V, true, true><<<g, t, b->evenLds, b->stream>>>(doms, c.even, c.evenRange, cod);
	else if (fl)   k_batch_even<Codec, V, true, false><<<g, t, b->evenLds, b->stream>>>(doms, c.even, c.evenRange, cod);
	else if (nt)   k_batch_even<Codec, V, false, true><<<g, t, b->evenLds, b->stream>>>(doms, c.even, c.evenRange, cod);
	else           k_batch_even<Codec, V, false, false><<<g, t, b->evenLds, b->stream>>>(doms, c.even, c.evenRange, cod);
}

template <class Codec, bool RLE>
static void launch_batch_odd(Batch * b, const ChunkRange & c, Codec cod, bool fl, bool fs)
{
	using R = typename Codec::R;
	const BatchDom<R> * doms = (const BatchDom<R> *)b->dDoms;
	const dim3 g(c.oddGrid), t(b->tune.oddBlock);
	const uint32_t gs = odd_group_shift(b);
	if (fs)      k_batch_odd<Codec, true, true, RLE><<<g, t, 0, b->stream>>>(doms, c.odd, c.oddRange, cod, gs);
	else if (fl) k_batch_odd<Codec, true, false, RLE><<<g, t, 0, b->stream>>>(doms, c.odd, c.oddRange, cod, gs);
	else         k_batch_odd<Codec, false, false, RLE><<<g, t, 0, b->stream>>>(doms, c.odd, c.oddRange, cod, gs);
}

template <class Codec>
static void launch_even_v(Batch * b, const ChunkRange & c, Codec cod, int v, bool fl)
{
	if constexpr (Codec::VMAX >= 4) { if (v == 4) { launch_batch_even<Codec, 4>(b, c, cod, fl); return; } }
	if constexpr (Codec::VMAX >= 2) { if (v == 2) { launch_batch_even<Codec, 2>(b, c, cod, fl); return; } }
	launch_batch_even<Codec, 1>(b, c, cod, fl);
}

template <class Codec>
static void launch_odd_rle(Batch * b, const ChunkRange & c, Codec cod, bool fl, bool fs)
{
	if (b->args.adjRle) launch_batch_odd<Codec, true>(b, c, cod, fl, fs);
	else launch_batch_odd<Codec, false>(b, c, cod, fl, fs);
}

// One even and one odd launch over the chunks of c.
template <class Codec>
static void launch_pair(Batch * b, const ChunkRange & c, Codec cod, int v, bool force, bool first, bool last)
{
	launch_even_v(b, c, cod, v, force && !first);
	launch_odd_rle(b, c, cod, force, force && last);
}

// `steps` steps over the chunks of c.
static void launch_range_steps(Batch * b, const ChunkRange & c, int steps)
{
	const int v = b->tableVec;
	const bool force = b->args.forceOn != 0;
	const float s16 = f16_scale(b);
	for (int it = 0; it < steps; it += 2) {
		const bool first = it == 0, last = it + 2 >= steps;
		switch (b->args.prec) {
			case PREC_FP32S: launch_pair(b, c, CodecF32{}, v, force, first, last); break;
			case PREC_FP16S: launch_pair(b, c, CodecF16{s16, 1.0f / s16}, v, force, first, last); break;
			default: launch_pair(b, c, CodecF64{}, v, force, first, last);
		}
	}
}

// `steps` steps of every active domain, launched directly: all chunks per
// launch, or group by group with every group taking all the steps before the
// next one.
static void launch_steps(Batch * b, int steps)
{
	for (size_t g = 0; g < slot_count(b); ++g) launch_range_steps(b, slot_chunks(b, g), steps);
}

static void drop_slot(Batch::Slot & s)
{
	if (s.exec) (void)hipGraphExecDestroy(s.exec);
	if (s.graph) (void)hipGraphDestroy(s.graph);
	s = Batch::Slot();
}

static void drop_graphs(Batch * b)
{
	for (Batch::Slot & s : b->slots) drop_slot(s);
	b->slots.clear();
}

// The launches of `steps` steps over slot g's range recorded as one graph
// (kernel nodes only).
static bool capture_slot(Batch * b, size_t g, int steps)
{
	Batch::Slot & s = b->slots[g];
	drop_slot(s);
	const ChunkRange c = slot_chunks(b, g);
	(void)hipGetLastError();
	if (!hip_ok(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return false;
	launch_range_steps(b, c, steps);
	const hipError_t launchErr = hipGetLastError();
	const hipError_t endErr = hipStreamEndCapture(b->stream, &s.graph);
	if (!hip_ok(launchErr, "batch step launch") || !hip_ok(endErr, "hipStreamEndCapture")) { drop_slot(s); return false; }
	if (!hip_ok(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0), "hipGraphInstantiate")) { drop_slot(s); return false; }
	s.range = c.evenRange;
	s.eGrid = c.evenGrid;
	s.oGrid = c.oddGrid;
	s.steps = steps;
	s.nt = c.nt;
	++b->captures;
	return true;
}

// Every slot's graph ready for `steps` steps over its current range: kept
// when it was captured for the same range words and step count with grids at
// least as large as the range needs now (a persistent grid larger than its
// chunk range only leaves blocks idle), captured again otherwise.
static bool ready_graphs(Batch * b, int steps)
{
	if (b->graphDirty) {
		drop_graphs(b);
		b->graphDirty = false;
	}
	const size_t n = slot_count(b);
	if (b->slots.size() < n) b->slots.resize(n);
	for (size_t g = 0; g < n; ++g) {
		const Batch::Slot & s = b->slots[g];
		const ChunkRange c = slot_chunks(b, g);
		if (s.exec && s.range == c.evenRange && s.steps == steps && s.eGrid >= c.evenGrid && s.oGrid >= c.oddGrid &&
		    s.nt == c.nt)
			continue;
		if (!capture_slot(b, g, steps)) return false;
	}
	return true;
}

template <class Codec>
static void launch_moments(Batch * b, Codec cod)
{
	using R = typename Codec::R;
	k_batch_moments<Codec><<<(unsigned)b->hMom.size(), 256, 0, b->stream>>>((const BatchDom<R> *)b->dDoms, b->dMom, cod, b->dPartial);
}

// FP64 mass and x momentum of every active domain: kernels and an
// asynchronous copy into pinned host memory, read by batch_finish.
static bool batch_moments_launch(Batch * b)
{
	const size_t nA = b->activeIds.size();
	if (nA == 0) return true;
	if (b->hSumsCap < 2 * nA) {
		host_free(b->hSumsPinned);
		b->hSumsCap = 0;
		b->hSumsPinned = (double *)host_alloc(sizeof(double) * 2 * nA);
		if (!b->hSumsPinned) return fail("allocating %zu bytes of pinned memory failed.", sizeof(double) * 2 * nA);
		b->hSumsCap = 2 * nA;
	}
	const float s16 = f16_scale(b);
	if (!b->hMom.empty()) {
		switch (b->args.prec) {
			case PREC_FP32S: launch_moments(b, CodecF32{}); break;
			case PREC_FP16S: launch_moments(b, CodecF16{s16, 1.0f / s16}); break;
			default: launch_moments(b, CodecF64{});
		}
	}
	k_batch_reduce<<<(unsigned)nA, 256, 0, b->stream>>>(b->dSeg, b->dPartial, b->dSums);
	return hip_ok(hipMemcpyAsync(b->hSumsPinned, b->dSums, sizeof(double) * 2 * nA, hipMemcpyDeviceToHost, b->stream),
	              "moment download");
}

// Starts one interval of `steps` for all active domains (graph replays, one
// per slot, then the moment kernels); returns without waiting.
static bool batch_launch(Batch * b, int steps)
{
	if (b->pendingSteps) return fail("LbmHipBatchLaunch: the previous interval has not been waited for.");
	steps += steps & 1;
	bool any = false;
	for (BatchEntry & x : b->e) {
		if (x.state != 1) continue;
		any = true;
		if (!list_upload(x.L)) return false;       // SetNode changes since the last interval
	}
	if (!any || steps <= 0) return true;
	if (b->tablesDirty && !rebuild_tables(b)) return false;
	if (!ready_graphs(b, steps)) return false;
	(void)hipGetLastError();
	if (!hip_ok(hipEventRecord(b->ev[0], b->stream), "hipEventRecord")) return false;
	b->lastSlots = slot_count(b);
	for (size_t g = 0; g < b->lastSlots; ++g) {
		if (!hip_ok(hipGraphLaunch(b->slots[g].exec, b->stream), "hipGraphLaunch")) return false;
		++b->replays;
	}
	if (!hip_ok(hipEventRecord(b->ev[1], b->stream), "hipEventRecord")) return false;
	if (!batch_moments_launch(b)) return false;
	if (!hip_ok(hipGetLastError(), "moment kernels")) return false;
	b->pendingSteps = steps;
	return true;
}

// Waits for the launched interval and applies the convergence test.
static bool batch_finish(Batch * b, double tol)
{
	const int steps = b->pendingSteps;
	if (!steps) return true;
	b->pendingSteps = 0;
	if (!hip_ok(hipStreamSynchronize(b->stream), "batch interval")) return false;
	float ms = 0.0f;
	if (!hip_ok(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]), "hipEventElapsedTime")) return false;
	b->gpuMs += ms;
	for (size_t i = 0; i < b->activeIds.size(); ++i) {
		BatchEntry & x = b->e[b->activeIds[i]];
		x.st.mass = b->hSumsPinned[2 * i] + (b->args.prec == PREC_FP64 ? 0.0 : (double)x.L->nFluid);
		x.st.momentumX = b->hSumsPinned[2 * i + 1];
		b->flups += (double)x.L->nFluid * steps;
		x.st.steps += steps;
		x.st.seconds += ms / 1e3;
		x.L->c.iteration = x.st.steps - 1;
		x.L->c.shadowStale = 1;
		const double jx = x.st.momentumX;
		if (x.havePrev) {
			const double d = fabs(jx - x.jxPrev), a = fabs(jx);
			x.st.relChange = a > 0.0 ? d / a : (d > 0.0 ? INFINITY : 0.0);
			if (x.st.relChange <= tol) x.st.converged = 1;
		}
		x.jxPrev = jx;
		x.havePrev = true;
		if (x.st.converged || (x.maxSteps > 0 && x.st.steps >= x.maxSteps)) {
			x.state = 2;
			x.st.active = 0;
			b->tablesDirty = true;
		}
	}
	return true;
}

static bool batch_step(Batch * b, int steps, double tol)
{
	return batch_launch(b, steps) && batch_finish(b, tol);
}

static void batch_free_tables(Batch * b)
{
	drop_graphs(b);
	(void)hipFree(b->dDoms);
	(void)hipFree(b->dEven);
	(void)hipFree(b->dOdd);
	(void)hipFree(b->dMom);
	(void)hipFree(b->dCount);
	(void)hipFree(b->dSeg);
	(void)hipFree(b->dPartial);
	(void)hipFree(b->dSums);
	(void)hipFree(b->dGroupRange);
}

static bool idle(const Batch * b, const char * who)
{
	if (!b) return fail("%s: invalid batch.", who);
	if (b->pendingSteps) return fail("%s: an interval is running; call LbmHipBatchWait first.", who);
	return true;
}

static BatchEntry * entry(Batch * b, int id, const char * who)
{
	if (!b || id < 0 || id >= (int)b->e.size() || b->e[id].state == 0) {
		fail("%s: no domain %d in this batch.", who, id);
		return nullptr;
	}
	return &b->e[id];
}

}  // namespace lbm

using namespace lbm;

extern "C" {

LbmHipBatch * LbmHipBatchCreate(const Parameters * params)
{
	Batch * b = new (std::nothrow) Batch();
	if (!b) { fail("allocating the batch failed."); return nullptr; }
	const int n = params ? params->nKernelArgs : 0;
	b->words.emplace_back("-shadow");        // host copies on demand (LbmHipBatchSync)
	b->words.emplace_back("lazy");
	for (int i = 0; i < n; ++i) b->words.emplace_back(params->KernelArgs[i]);
	b->words.emplace_back("-quiet");
	for (std::string & w : b->words) b->argv.push_back(&w[0]);
	b->params.nKernelArgs = (int)b->argv.size();
	b->params.KernelArgs = b->argv.data();
	bool ok = parse_kernel_args(&b->params, &b->args, true);
	int nDev = 0;
	if (ok) {
		const hipError_t err = hipGetDeviceCount(&nDev);
		if (err == hipErrorNoDevice || (err == hipSuccess && nDev < 1))
			ok = fail("no HIP device found; the list/aa kernels of this library run on the GPU only.");
		else ok = hip_ok(err, "hipGetDeviceCount");
	}
	ok = ok && hip_ok(hipGetDevice(&b->device), "hipGetDevice")
	        && hip_ok(hipDeviceGetAttribute(&b->nCU, hipDeviceAttributeMultiprocessorCount, b->device), "hipDeviceGetAttribute")
	        && hip_ok(hipDeviceGetAttribute(&b->ldsPerCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, b->device),
	                  "hipDeviceGetAttribute")
	        && hip_ok(hipDeviceGetAttribute(&b->ldsPerBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device),
	                  "hipDeviceGetAttribute")
	        && hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate")
	        && hip_ok(hipEventCreate(&b->ev[0]), "hipEventCreate") && hip_ok(hipEventCreate(&b->ev[1]), "hipEventCreate");
	if (!ok) {
		LbmHipBatchDestroy(reinterpret_cast<LbmHipBatch *>(b));
		return nullptr;
	}
	// defaults: the single-lattice launch shapes
	b->tune.evenBlock = b->args.evenCap ? 512 : b->args.blockSize;
	b->tune.evenCap = b->args.evenCap;
	b->tune.vec = b->args.evenVec ? vec_max(b->args.prec) : 1;
	b->tune.oddBlock = b->args.blockSize;
	b->tune.groupMB = kDefaultGroupMB;
	b->tune.oddGroup = (int)kXcdGroup;
	// even-step accesses per launch range by what it touches (-1), unless
	// -nt on|off forces them: a range that fits the last-level cache streams
	// faster with plain accesses (512 porous samples, 256 MiB groups: FP32
	// 408 against 469 ms, FP64 687 against 762 ms per 100 steps)
	b->tune.nontemporal = b->args.nontemporal;
	return reinterpret_cast<LbmHipBatch *>(b);
}

void LbmHipBatchDestroy(LbmHipBatch * handle)
{
	Batch * b = as_batch(handle);
	if (!b) return;
	for (BatchEntry & x : b->e)
		if (x.kd) D3Q19ListAaPvDeinit_PushSoAHip(nullptr, &x.kd);
	if (b->pendingSteps) (void)hipStreamSynchronize(b->stream);
	batch_free_tables(b);
	host_free(b->hSumsPinned);
	if (b->ev[0]) (void)hipEventDestroy(b->ev[0]);
	if (b->ev[1]) (void)hipEventDestroy(b->ev[1]);
	if (b->stream) (void)hipStreamDestroy(b->stream);
	delete b;
}

int LbmHipBatchAdd(LbmHipBatch * handle, LatticeDesc * ld, double omega, double xForce, int maxSteps)
{
	// allowed while an interval runs: the new domain joins at the next one
	Batch * b = as_batch(handle);
	if (!b || !ld) { fail("LbmHipBatchAdd: invalid arguments."); return -1; }
	if (!(omega > 0.0 && omega < 2.0)) { fail("omega must be in (0, 2), got %f.", omega); return -1; }
	KernelData * kd = nullptr;
	D3Q19ListAaPvInit_PushSoAHip(ld, &kd, &b->params);
	if (!kd) return -1;
	ListData * L = list_data(kd);
	// f = w (rho 1, u 0) on the device: zero bits for shifted storage (as
	// Init leaves them), the weights for FP64
	if (b->args.prec == PREC_FP64 && L->nFluid > 0) {
		const uint64_t total = 19ull * (uint64_t)L->nFluid;
		(void)hipGetLastError();
		k_fill_weights<<<(unsigned)((total + 255) / 256), 256, 0, L->c.stream>>>((double *)L->dPdf, (uint32_t)L->nFluid,
		                                                                         (uint32_t)L->nCells);
		if (!hip_ok(hipGetLastError(), "weight fill") || !hip_ok(hipStreamSynchronize(L->c.stream), "weight fill")) {
			D3Q19ListAaPvDeinit_PushSoAHip(nullptr, &kd);
			return -1;
		}
	}
	L->c.shadowDirty = 0;
	L->c.shadowStale = 1;          // the device holds the state; the host copy follows on LbmHipBatchSync
	int id = -1;
	for (size_t i = 0; i < b->e.size(); ++i)
		if (b->e[i].state == 0) { id = (int)i; break; }
	if (id < 0) { id = (int)b->e.size(); b->e.emplace_back(); }
	BatchEntry & x = b->e[id];
	x = BatchEntry();
	x.kd = kd;
	x.L = L;
	x.state = 1;
	x.omega = omega;
	x.xForce = xForce;
	x.maxSteps = maxSteps;
	x.st.active = 1;
	x.st.relChange = -1.0;
	x.st.nFluid = L->nFluid;
	b->tablesDirty = true;
	return id;
}

int LbmHipBatchRemove(LbmHipBatch * handle, int id)
{
	Batch * b = as_batch(handle);
	if (!idle(b, "LbmHipBatchRemove")) return -1;
	BatchEntry * x = entry(b, id, "LbmHipBatchRemove");
	if (!x) return -1;
	D3Q19ListAaPvDeinit_PushSoAHip(nullptr, &x->kd);
	*x = BatchEntry();
	b->tablesDirty = true;
	return 0;
}

KernelData * LbmHipBatchKernelData(LbmHipBatch * handle, int id)
{
	BatchEntry * x = entry(as_batch(handle), id, "LbmHipBatchKernelData");
	return x ? x->kd : nullptr;
}

int LbmHipBatchStatusOf(LbmHipBatch * handle, int id, LbmHipBatchStatus * out)
{
	BatchEntry * x = entry(as_batch(handle), id, "LbmHipBatchStatusOf");
	if (!x || !out) return -1;
	*out = x->st;
	return 0;
}

int LbmHipBatchStep(LbmHipBatch * handle, int steps, double tol, int * nActive)
{
	Batch * b = as_batch(handle);
	if (!b || steps < 0 || !(tol >= 0.0)) { fail("LbmHipBatchStep: invalid arguments."); return -1; }
	if (!batch_step(b, steps, tol)) return -1;
	if (nActive) {
		int k = 0;
		for (const BatchEntry & x : b->e) k += x.state == 1;
		*nActive = k;
	}
	return 0;
}

int LbmHipBatchLaunch(LbmHipBatch * handle, int steps)
{
	Batch * b = as_batch(handle);
	if (!b || steps < 0) { fail("LbmHipBatchLaunch: invalid arguments."); return -1; }
	return batch_launch(b, steps) ? 0 : -1;
}

int LbmHipBatchWait(LbmHipBatch * handle, double tol, int * nActive)
{
	Batch * b = as_batch(handle);
	if (!b || !(tol >= 0.0)) { fail("LbmHipBatchWait: invalid arguments."); return -1; }
	if (!batch_finish(b, tol)) return -1;
	if (nActive) {
		int k = 0;
		for (const BatchEntry & x : b->e) k += x.state == 1;
		*nActive = k;
	}
	return 0;
}

int LbmHipBatchRun(LbmHipBatch * handle, int checkEvery, double tol)
{
	Batch * b = as_batch(handle);
	if (!b || checkEvery <= 0) { fail("LbmHipBatchRun: invalid arguments."); return -1; }
	for (const BatchEntry & x : b->e)
		if (x.state == 1 && x.maxSteps <= 0 && !(tol > 0.0)) {
			fail("LbmHipBatchRun: a domain without maxSteps needs tol > 0 to end.");
			return -1;
		}
	int active = 1;
	while (active > 0)
		if (LbmHipBatchStep(handle, checkEvery, tol, &active) != 0) return -1;
	return 0;
}

int LbmHipBatchSync(LbmHipBatch * handle, int id)
{
	Batch * b = as_batch(handle);
	if (!idle(b, "LbmHipBatchSync")) return -1;
	if (id >= 0) {
		BatchEntry * x = entry(b, id, "LbmHipBatchSync");
		return (x && (!x->L->c.shadowStale || list_download(x->L))) ? 0 : -1;
	}
	if (!b) { fail("LbmHipBatchSync: invalid batch."); return -1; }
	for (BatchEntry & x : b->e)
		if (x.state != 0 && x.L->c.shadowStale && !list_download(x.L)) return -1;
	return 0;
}

int LbmHipBatchGetTuning(LbmHipBatch * handle, LbmHipBatchTuning * out)
{
	Batch * b = as_batch(handle);
	if (!b || !out) { fail("LbmHipBatchGetTuning: invalid arguments."); return -1; }
	*out = b->tune;
	out->vec = batch_vec(b);
	return 0;
}

int LbmHipBatchSetTuning(LbmHipBatch * handle, const LbmHipBatchTuning * t)
{
	Batch * b = as_batch(handle);
	if (!b || !t) { fail("LbmHipBatchSetTuning: invalid arguments."); return -1; }
	if (!idle(b, "LbmHipBatchSetTuning")) return -1;
	const bool okBlock = t->evenBlock >= 64 && t->evenBlock <= 512 && t->evenBlock % 64 == 0 &&
	                     t->oddBlock >= 64 && t->oddBlock <= 512 && t->oddBlock % 64 == 0;
	if (!okBlock || (t->vec != 1 && t->vec != 2 && t->vec != 4) || t->vec > vec_max(b->args.prec) ||
	    t->groupMB < 0 || t->groupMB > (1 << 20) || t->oddGroup < 0 || t->oddGroup > 256 ||
	    (t->oddGroup & (t->oddGroup - 1))) {
		fail("LbmHipBatchSetTuning: blocks must be 64..512 in steps of 64, vec 1, 2 (fp32s) or 4 (fp16s), "
		     "groupMB 0..2^20, oddGroup 0 or a power of two up to 256.");
		return -1;
	}
	b->tune = *t;
	if (b->tune.oddGroup == 0) b->tune.oddGroup = (int)kXcdGroup;
	b->tune.evenCap = t->evenCap != 0;
	b->tune.nontemporal = t->nontemporal < 0 ? -1 : t->nontemporal != 0;
	b->tablesDirty = true;
	b->graphDirty = true;
	return 0;
}

int LbmHipBatchStats(LbmHipBatch * handle, double * gpuSeconds, double * flups)
{
	Batch * b = as_batch(handle);
	if (!b) { fail("LbmHipBatchStats: invalid batch."); return -1; }
	if (gpuSeconds) *gpuSeconds = b->gpuMs / 1e3;
	if (flups) *flups = b->flups;
	return 0;
}

int LbmHipBatchGraphStats(LbmHipBatch * handle, int * slots, long long * captures, long long * replays)
{
	Batch * b = as_batch(handle);
	if (!b) { fail("LbmHipBatchGraphStats: invalid batch."); return -1; }
	if (slots) *slots = (int)b->lastSlots;
	if (captures) *captures = b->captures;
	if (replays) *replays = b->replays;
	return 0;
}

// Times `pairs` step pairs of every candidate launch shape on the active
// domains and keeps the fastest even and odd shapes.  The lattices are
// stepped while tuning, so each is copied aside on the device first and
// copied back afterwards.
int LbmHipBatchAutotune(LbmHipBatch * handle, int pairs, LbmHipBatchTuning * out)
{
	Batch * b = as_batch(handle);
	if (!b || pairs <= 0) { fail("LbmHipBatchAutotune: invalid arguments."); return -1; }
	if (!idle(b, "LbmHipBatchAutotune")) return -1;
	std::vector<void *> backup(b->e.size(), nullptr);
	auto release_backup = [&]() { for (void * p : backup) (void)hipFree(p); };
	for (size_t i = 0; i < b->e.size(); ++i) {
		BatchEntry & x = b->e[i];
		if (x.state != 1) continue;
		if (!list_upload(x.L)) { release_backup(); return -1; }
		const size_t bytes = elem_bytes(b->args.prec) * 19 * (size_t)x.L->nCells;
		if (!hip_ok(hipMalloc(&backup[i], bytes), "allocating the tuning backup") ||
		    !hip_ok(hipMemcpyAsync(backup[i], x.L->dPdf, bytes, hipMemcpyDeviceToDevice, b->stream), "tuning backup")) {
			release_backup();
			return -1;
		}
	}
	const LbmHipBatchTuning saved = b->tune;
	const int vm = b->args.evenVec ? vec_max(b->args.prec) : 1;
	struct Cand { int block, cap, vec; };
	std::vector<Cand> evens;
	for (int cap = 1; cap >= 0; --cap)
		for (int blk : {256, 512})
			for (int v = vm; v >= 1; v /= 2) evens.push_back({blk, cap, v});
	struct OddCand { int block, group; };
	std::vector<OddCand> odds;
	for (int o : {128, 256, 512})
		for (int g : {8, 32, 128}) odds.push_back({o, g});
	float bestEven = 1e30f, bestOdd = 1e30f;
	LbmHipBatchTuning best = saved;
	bool ok = true;
	hipEvent_t t0 = b->ev[0], t1 = b->ev[1];
	// evenOnly / oddOnly: `pairs` launches of that step over all chunks;
	// neither: `pairs` step pairs through launch_steps (grouped or not)
	auto time_it = [&](bool evenOnly, bool oddOnly, float * ms) -> bool {
		if (b->tablesDirty && !rebuild_tables(b)) return false;
		(void)hipGetLastError();
		const float s16 = f16_scale(b);
		const int v = b->tableVec;
		const ChunkRange all = all_chunks(b);
		if (!hip_ok(hipEventRecord(t0, b->stream), "hipEventRecord")) return false;
		if (!evenOnly && !oddOnly) launch_steps(b, 2 * pairs);
		for (int p = 0; p < pairs && (evenOnly || oddOnly); ++p) {
			if (evenOnly) {
				switch (b->args.prec) {
					case PREC_FP32S: launch_even_v(b, all, CodecF32{}, v, false); break;
					case PREC_FP16S: launch_even_v(b, all, CodecF16{s16, 1.0f / s16}, v, false); break;
					default: launch_even_v(b, all, CodecF64{}, v, false);
				}
			} else {
				switch (b->args.prec) {
					case PREC_FP32S: launch_odd_rle(b, all, CodecF32{}, false, false); break;
					case PREC_FP16S: launch_odd_rle(b, all, CodecF16{s16, 1.0f / s16}, false, false); break;
					default: launch_odd_rle(b, all, CodecF64{}, false, false);
				}
			}
		}
		if (!hip_ok(hipGetLastError(), "tuning launch")) return false;
		if (!hip_ok(hipEventRecord(t1, b->stream), "hipEventRecord")) return false;
		if (!hip_ok(hipEventSynchronize(t1), "tuning")) return false;
		return hip_ok(hipEventElapsedTime(ms, t0, t1), "hipEventElapsedTime");
	};
	LbmHipBatchTuning base = saved;
	base.groupMB = 0;
	for (const Cand & c : evens) {
		b->tune = base;
		b->tune.evenBlock = c.block;
		b->tune.evenCap = c.cap;
		b->tune.vec = c.vec;
		b->tablesDirty = true;
		b->graphDirty = true;
		float ms = 0.0f;
		if (!(ok = time_it(true, false, &ms))) break;
		if (ms < bestEven) { bestEven = ms; best.evenBlock = c.block; best.evenCap = c.cap; best.vec = c.vec; }
	}
	for (const OddCand & o : odds) {
		if (!ok) break;
		b->tune = base;
		b->tune.oddBlock = o.block;
		b->tune.oddGroup = o.group;
		b->tablesDirty = true;
		b->graphDirty = true;
		float ms = 0.0f;
		if (!(ok = time_it(false, true, &ms))) break;
		if (ms < bestOdd) { bestOdd = ms; best.oddBlock = o.block; best.oddGroup = o.group; }
	}
	// cache-sized groups with the chosen shapes: whole pairs, group by group,
	// each group size with plain and with non-temporal even-step accesses
	// (a group that fits the last-level cache streams faster with plain
	// ones: on the 512-sample FP32 batch 512 MiB groups took 384 ms per 100
	// steps with plain and 435 ms with non-temporal accesses)
	float bestPair = 1e30f;
	best.groupMB = 0;
	for (int mb : {0, 64, 128, 192, 256, 384, 512, 768, 1024}) {
		for (int nt = 0; nt <= 1 && ok; ++nt) {
			b->tune = best;
			b->tune.groupMB = mb;
			b->tune.nontemporal = nt;
			b->tablesDirty = true;
			b->graphDirty = true;
			float ms = 0.0f;
			if (!(ok = time_it(false, false, &ms))) break;
			if (ms < bestPair) { bestPair = ms; best.groupMB = mb; best.nontemporal = nt; }
		}
		if (!ok) break;
	}
	// with groups the data sits in the cache: the launch shapes again, as whole pairs
	if (ok && best.groupMB > 0) {
		const LbmHipBatchTuning grouped = best;
		for (const Cand & c : evens) {
			b->tune = grouped;
			b->tune.evenBlock = c.block;
			b->tune.evenCap = c.cap;
			b->tune.vec = c.vec;
			b->tablesDirty = true;
			float ms = 0.0f;
			if (!(ok = time_it(false, false, &ms))) break;
			if (ms < bestPair) { bestPair = ms; best.evenBlock = c.block; best.evenCap = c.cap; best.vec = c.vec; }
		}
		for (const OddCand & o : odds) {
			if (!ok) break;
			b->tune = best;
			b->tune.oddBlock = o.block;
			b->tune.oddGroup = o.group;
			b->tablesDirty = true;
			float ms = 0.0f;
			if (!(ok = time_it(false, false, &ms))) break;
			if (ms < bestPair) { bestPair = ms; best.oddBlock = o.block; best.oddGroup = o.group; }
		}
		// non-temporal even-step accesses stream past the cache the groups rely on
		for (int nt = 0; nt <= 1 && ok; ++nt) {
			b->tune = best;
			b->tune.nontemporal = nt;
			b->tablesDirty = true;
			float ms = 0.0f;
			if (!(ok = time_it(false, false, &ms))) break;
			if (ms < bestPair) { bestPair = ms; best.nontemporal = nt; }
		}
	}
	b->tune = ok ? best : saved;
	b->tune.evenMs = ok ? bestEven / pairs : 0.0;
	b->tune.oddMs = ok ? bestOdd / pairs : 0.0;
	b->tune.pairMs = ok ? bestPair / pairs : 0.0;
	b->tablesDirty = true;
	b->graphDirty = true;
	// restore the lattices
	for (size_t i = 0; i < b->e.size(); ++i) {
		if (!backup[i]) continue;
		const size_t bytes = elem_bytes(b->args.prec) * 19 * (size_t)b->e[i].L->nCells;
		if (!hip_ok(hipMemcpyAsync(b->e[i].L->dPdf, backup[i], bytes, hipMemcpyDeviceToDevice, b->stream), "tuning restore"))
			ok = false;
	}
	if (!hip_ok(hipStreamSynchronize(b->stream), "tuning restore")) ok = false;
	release_backup();
	if (out) *out = b->tune;
	return ok ? 0 : -1;
}

}  // extern "C"
