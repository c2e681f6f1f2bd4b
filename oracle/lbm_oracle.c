/*
 * lbm_oracle.c -- CPU restatement of the reference's D3Q19 TRT AA path.
 *
 * TEST INFRASTRUCTURE ONLY (see lbm_oracle.h): the checker for the HIP
 * implementation.  Nothing in the product links or loads this file.
 *
 * Built with -O2 -ffp-contract=off and without -ffast-math, so every
 * expression below is evaluated in the written order with one IEEE-754
 * rounding per operation, which is how the reference's scalar C evaluates it
 * (gcc -O3 -mavx -std=c99 has no FMA instructions to contract into).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <limits.h>

/* Stencil, src/Kernel.h:121-140: N S E W NE SE NW SW T TN TE TW TS B BS BN BW BE C */
const int ORC_CX[ORC_NDIR]  = { 0, 0, 1,-1, 1, 1,-1,-1, 0, 0, 1,-1, 0, 0, 0, 0,-1, 1, 0 };
const int ORC_CY[ORC_NDIR]  = { 1,-1, 0, 0, 1,-1, 1,-1, 0, 1, 0, 0,-1, 0,-1, 1, 0, 0, 0 };
const int ORC_CZ[ORC_NDIR]  = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1,-1,-1,-1,-1,-1, 0 };
const int ORC_INV[ORC_NDIR] = { 1, 0, 3, 2, 7, 6, 5, 4,13,14,16,17,15, 8, 9,12,10,11,18 };
/* src/Kernel.c:769-772 */
#define W0 (1.0 / 3.0)
#define W1 (1.0 / 18.0)
#define W2 (1.0 / 36.0)
const double ORC_W[ORC_NDIR] = { W1, W1, W1, W1, W2, W2, W2, W2, W1, W2, W2, W2, W2,
                                 W1, W2, W2, W2, W2, W0 };

enum { N_ = 0, S_, E_, W_, NE_, SE_, NW_, SW_, T_, TN_, TE_, TW_, TS_, B_, BS_, BN_, BW_, BE_, C_ };

static inline int lidx(const int d[3], int x, int y, int z)
{
	return z * d[0] * d[1] + y * d[0] + x;          /* src/Lattice.h:75 */
}

/* ------------------------------------------------------------------------
 * Padding -- src/Padding.c:19-244
 * --------------------------------------------------------------------- */
int orc_pad_parse(const char * s, orc_pad_t * p)
{
	memset(p, 0, sizeof(*p));
	if (s == NULL) return -1;
	if (!strncmp("auto", s, 4)) {                   /* src/Padding.c:24-42 */
		p->nEntries = 2;
		p->modulus[0] = 2 * 1024 * 1024 * 8; p->offset[0] = 2 * 1024 * 1024 + 512;
		p->modulus[1] = 256 * 64;            p->offset[1] = 128;
		return 0;
	}
	if (!strncmp("no", s, 2) || !strncmp("off", s, 3)) return 0;

	int nCommas = 0;
	for (const char * t = s; *t; ++t) if (*t == ',') ++nCommas;
	if (nCommas + 1 > 16) return -10;

	int idx = 0, seen = 0, modSeen = 0, offSeen = 0;
	const char * t = s;
	while (*t) {                                    /* src/Padding.c:84-176 */
		if (*t == ',') {
			if (!modSeen) return -2;
			modSeen = 0; offSeen = 0; ++idx; ++t;
		}
		else if (*t == '+') {
			if (!modSeen) return -3;
			if (offSeen)  return -4;
			++t;
		}
		else if (*t >= '0' && *t <= '9') {
			char * end;
			errno = 0;
			long v = strtol(t, &end, 10);
			if (errno != 0 || v > INT_MAX) return -5;
			if (!modSeen) { p->modulus[idx] = (int)v; modSeen = 1; ++seen; }
			else {
				if (offSeen) return -4;
				p->offset[idx] = (int)v; offSeen = 1;
			}
			t = end;
		}
		else return -6;
	}
	if (seen != nCommas + 1) return -7;
	for (int i = 0; i < nCommas + 1; ++i)
		if (p->offset[i] >= p->modulus[i]) return -8;
	p->nEntries = seen;
	return 0;
}

int orc_pad_cells(int nCells, int cellBytes, const orc_pad_t * pad)
{
	orc_pad_t autoPad;
	if (pad == NULL) { orc_pad_parse("auto", &autoPad); pad = &autoPad; }
	for (int i = 0; i < pad->nEntries; ++i) {       /* src/Padding.c:215-241 */
		int add;
		if (pad->modulus[i] == 0) {
			add = pad->offset[i] * cellBytes;       /* reference multiplies by the cell size here */
		} else {
			int m = pad->modulus[i] / cellBytes;
			if (m == 0) return -1;
			int o = pad->offset[i] / cellBytes;
			add = (o + m - nCells % m) % m;
		}
		nCells += add;
	}
	return nCells;
}

/* ------------------------------------------------------------------------
 * Geometry -- src/Geometry.c:90-281
 * --------------------------------------------------------------------- */
int orc_geometry(int type, int blockSize, const int dims[3], const int periodicIn[3],
                 int * m, int pf[3])
{
	const int X = dims[0], Y = dims[1], Z = dims[2];
	pf[0] = periodicIn[0]; pf[1] = periodicIn[1]; pf[2] = periodicIn[2];
	for (int i = 0; i < X * Y * Z; ++i) m[i] = 1;

	if (type == ORC_GEO_CHANNEL || type == ORC_GEO_BLOCKS || type == ORC_GEO_PIPE) pf[0] = 1;
	else if (type == ORC_GEO_FLUID) { pf[0] = pf[1] = pf[2] = 1; }

	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) {
		m[lidx(dims, 0, y, z)] = pf[0] ? 1 : 0;
		m[lidx(dims, X - 1, y, z)] = pf[0] ? 1 : 0;
	}
	for (int z = 0; z < Z; ++z) for (int x = 0; x < X; ++x) {
		m[lidx(dims, x, 0, z)] = pf[1] ? 1 : 0;
		m[lidx(dims, x, Y - 1, z)] = pf[1] ? 1 : 0;
	}
	for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		m[lidx(dims, x, y, 0)] = pf[2] ? 1 : 0;
		m[lidx(dims, x, y, Z - 1)] = pf[2] ? 1 : 0;
	}

	if (type == ORC_GEO_PIPE) {
		double cz = Z / 2.0 - 0.5, cy = Y / 2.0 - 0.5;
		double dmin = (Y < Z) ? Y : Z;
		double r2 = (dmin / 2 - 1) * (dmin / 2 - 1);
		for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y)
			if ((z - cz) * (z - cz) + (y - cy) * (y - cy) >= r2)
				for (int x = 0; x < X; ++x) m[lidx(dims, x, y, z)] = 0;
	}
	else if (type == ORC_GEO_BLOCKS && blockSize > 0) {
		int b = blockSize;
		for (int z = 0; z < Z; ++z) {
			if ((z % (2 * b)) < b) continue;
			for (int y = 0; y < Y; ++y) {
				if ((y % (2 * b)) < b) continue;
				for (int x = 0; x < X; ++x)
					if ((x % (2 * b)) >= b) m[lidx(dims, x, y, z)] = 0;
			}
		}
	}
	int nFluid = 0;
	for (int i = 0; i < X * Y * Z; ++i) nFluid += (m[i] != 0);
	return nFluid;
}

/* ------------------------------------------------------------------------
 * List build -- src/BenchKernelD3Q19ListAaPvCommon.c:627-793
 * --------------------------------------------------------------------- */
int orc_list_number(const int * mask, const int dims[3], const int blkIn[3],
                    uint32_t * grid, uint32_t * coords)
{
	const int X = dims[0], Y = dims[1], Z = dims[2];
	int b[3] = { blkIn[0] ? blkIn[0] : X, blkIn[1] ? blkIn[1] : Y, blkIn[2] ? blkIn[2] : Z };
	for (int i = 0; i < X * Y * Z; ++i) grid[i] = 0xffffffffu;
	uint32_t n = 0;
	for (int bx = 0; bx < X; bx += b[0])
	for (int by = 0; by < Y; by += b[1])
	for (int bz = 0; bz < Z; bz += b[2]) {
		int ex = bx + b[0] < X ? bx + b[0] : X;
		int ey = by + b[1] < Y ? by + b[1] : Y;
		int ez = bz + b[2] < Z ? bz + b[2] : Z;
		for (int x = bx; x < ex; ++x)
		for (int y = by; y < ey; ++y)
		for (int z = bz; z < ez; ++z) {
			int l = lidx(dims, x, y, z);
			if (mask[l] != 0) {
				grid[l] = n;
				coords[3 * n + 0] = x; coords[3 * n + 1] = y; coords[3 * n + 2] = z;
				++n;
			}
		}
	}
	return (int)n;
}

void orc_list_adjacency(const int * mask, const int dims[3], const int per[3],
                        const uint32_t * grid, const uint32_t * coords,
                        int nFluid, int nCells, uint32_t * adj)
{
	const int X = dims[0], Y = dims[1], Z = dims[2];
	for (int n = 0; n < nFluid; ++n) {
		int x = coords[3 * n], y = coords[3 * n + 1], z = coords[3 * n + 2];
		for (int d = 0; d < ORC_NLINK; ++d) {
			int nx = x + ORC_CX[d], ny = y + ORC_CY[d], nz = z + ORC_CZ[d];
			uint32_t v;
			int outX = nx < 0 || nx >= X, outY = ny < 0 || ny >= Y, outZ = nz < 0 || nz >= Z;
			if ((outX && per[0]) || (outY && per[1]) || (outZ && per[2])) {
				/* once one periodic coordinate leaves the box, every
				 * out-of-range coordinate is wrapped (:729-761) */
				int px = nx < 0 ? X - 1 : (nx >= X ? 0 : nx);
				int py = ny < 0 ? Y - 1 : (ny >= Y ? 0 : ny);
				int pz = nz < 0 ? Z - 1 : (nz >= Z ? 0 : nz);
				int l = lidx(dims, px, py, pz);
				if (mask[l] == 0) v = (uint32_t)ORC_INV[d] * nCells + n;
				else              v = (uint32_t)d * nCells + grid[l];
			}
			else if (outX || outY || outZ) {
				v = (uint32_t)ORC_INV[d] * nCells + n;
			}
			else {
				int l = lidx(dims, nx, ny, nz);
				if (mask[l] == 0) v = (uint32_t)ORC_INV[d] * nCells + n;
				else              v = (uint32_t)d * nCells + grid[l];
			}
			adj[(size_t)n * ORC_NLINK + d] = v;
		}
	}
}

/* Node ranges of the reference's OpenMP threads (:342-349). */
static int thread_start(int t, int nFluid, int nThreads)
{
	int per = nFluid / nThreads, rem = nFluid % nThreads;
	return t * per + (t < rem ? t : rem);
}

int orc_list_runs(const uint32_t * adj, int nFluid, int nThreads, uint32_t * runLen)
{
	if (nFluid <= 0) return 0;
	int nextBreakThread = 1;
	int r = 0;
	if (runLen) runLen[0] = 1;
	for (int n = 1; n < nFluid; ++n) {
		int diff = 0;
		for (int d = 0; d < ORC_NLINK; ++d)
			if (adj[(size_t)n * ORC_NLINK + d] != adj[(size_t)(n - 1) * ORC_NLINK + d] + 1) { diff = 1; break; }
		if (nextBreakThread < nThreads && thread_start(nextBreakThread, nFluid, nThreads) == n) {
			++nextBreakThread; diff = 1;
		}
		if (diff) { ++r; if (runLen) runLen[r] = 1; }
		else if (runLen) runLen[r] += 1;
	}
	return r + 1;
}

/* ------------------------------------------------------------------------
 * TRT collision -- src/BenchKernelD3Q19ListAaPv.c:380-465 (scalar tree),
 * :242-378 (AVX tree).  Identical formulas in src/BenchKernelD3Q19Aa.c:205-284.
 * --------------------------------------------------------------------- */
double orc_omega_odd(double omega, double magic)
{
	return 1.0 / (0.5 + magic / (1.0 / omega - 0.5));
}

void orc_trt_collide(const double p[ORC_NDIR], double q[ORC_NDIR],
                     double oe, double oo, int tree)
{
	const double w_0 = 1.0 / 3.0, w_1 = 1.0 / 18.0, w_2 = 1.0 / 36.0;
	const double w_1_x3 = w_1 * 3.0, w_1_nh = w_1 * 9.0 / 2.0;
	const double w_2_x3 = w_2 * 3.0, w_2_nh = w_2 * 9.0 / 2.0;
	double ux, uy, uz, dens, dir;

	if (tree == ORC_TREE_SCALAR) {
		ux = p[E_] + p[NE_] + p[SE_] + p[TE_] + p[BE_] - p[W_] - p[NW_] - p[SW_] - p[TW_] - p[BW_];
		uy = p[N_] + p[NE_] + p[NW_] + p[TN_] + p[BN_] - p[S_] - p[SE_] - p[SW_] - p[TS_] - p[BS_];
		uz = p[T_] + p[TE_] + p[TW_] + p[TN_] + p[TS_] - p[B_] - p[BE_] - p[BW_] - p[BN_] - p[BS_];
		dens = p[C_] + p[N_] + p[E_] + p[S_] + p[W_] + p[NE_] + p[SE_] + p[SW_] + p[NW_]
		     + p[T_] + p[TN_] + p[TE_] + p[TS_] + p[TW_] + p[B_] + p[BN_] + p[BE_] + p[BS_] + p[BW_];
		dir = dens - (ux * ux + uy * uy + uz * uz) * 3.0 / 2.0;
	} else {
		ux = (p[E_] + (p[NE_] + p[SE_])) + (p[TE_] + p[BE_]);
		ux = ux - p[W_]; ux = ux - p[NW_]; ux = ux - p[SW_]; ux = ux - p[TW_]; ux = ux - p[BW_];
		uy = (p[N_] + (p[NE_] + p[NW_])) + (p[TN_] + p[BN_]);
		uy = uy - p[S_]; uy = uy - p[SE_]; uy = uy - p[SW_]; uy = uy - p[TS_]; uy = uy - p[BS_];
		uz = (p[T_] + (p[TE_] + p[TW_])) + (p[TN_] + p[TS_]);
		uz = uz - p[B_]; uz = uz - p[BE_]; uz = uz - p[BW_]; uz = uz - p[BN_]; uz = uz - p[BS_];
		dens = p[C_] + (p[N_] + p[E_]);
		dens = dens + (p[S_] + p[W_]);
		dens = dens + (p[NE_] + p[SE_]);
		dens = dens + (p[SW_] + p[NW_]);
		dens = dens + (p[T_] + p[TN_]);
		dens = dens + (p[TE_] + p[TS_]);
		dens = dens + (p[TW_] + p[B_]);
		dens = dens + (p[BN_] + p[BE_]);
		dens = dens + (p[BS_] + p[BW_]);
		dir = dens - ((ux * ux + uy * uy) + uz * uz) * 1.5;
	}

	q[C_] = p[C_] - oe * (p[C_] - w_0 * dir);

	const double w1i = w_1 * dir, w2i = w_2 * dir;
	double ev, od, ui;
#define PAIR(a, b, uiexpr, wi, wnh, wx3)                                         \
	ui = (uiexpr);                                                               \
	if (tree == ORC_TREE_SCALAR) ev = oe * (0.5 * (p[a] + p[b]) - ui * ui * wnh - wi); \
	else                         ev = oe * (0.5 * (p[a] + p[b]) - ui * (ui * wnh) - wi); \
	od = oo * (0.5 * (p[a] - p[b]) - ui * wx3);                                  \
	q[a] = p[a] - ev - od;                                                       \
	q[b] = p[b] - ev + od;

	PAIR(N_,  S_,  uy,        w1i, w_1_nh, w_1_x3)
	PAIR(E_,  W_,  ux,        w1i, w_1_nh, w_1_x3)
	PAIR(T_,  B_,  uz,        w1i, w_1_nh, w_1_x3)
	PAIR(NW_, SE_, -ux + uy,  w2i, w_2_nh, w_2_x3)
	PAIR(NE_, SW_, ux + uy,   w2i, w_2_nh, w_2_x3)
	PAIR(TW_, BE_, -ux + uz,  w2i, w_2_nh, w_2_x3)
	PAIR(TE_, BW_, ux + uz,   w2i, w_2_nh, w_2_x3)
	PAIR(TS_, BN_, -uy + uz,  w2i, w_2_nh, w_2_x3)
	PAIR(TN_, BS_, uy + uz,   w2i, w_2_nh, w_2_x3)
#undef PAIR
}

/* ------------------------------------------------------------------------
 * Body force -- src/Kernel.c:759-803 (Luo forcing in node view)
 * --------------------------------------------------------------------- */
double orc_force_term(int d, double F)
{
	return 3.0 * ORC_W[d] * ORC_CX[d] * F;
}

/* ------------------------------------------------------------------------
 * List AA -- src/BenchKernelD3Q19ListAaPv.c:45-851,
 *            node view src/BenchKernelD3Q19ListAaPvCommon.c:127-223
 * --------------------------------------------------------------------- */
static void list_node_slots(const orc_list_t * L, uint32_t n, int parityEven, size_t slot[ORC_NDIR])
{
	const size_t nc = (size_t)L->nCells;
	slot[C_] = (size_t)C_ * nc + n;
	for (int d = 0; d < ORC_NLINK; ++d) {
		if (parityEven) slot[d] = L->adj[(size_t)n * ORC_NLINK + ORC_INV[d]];
		else            slot[d] = (size_t)d * nc + n;
	}
}

void orc_list_get_node(const orc_list_t * L, int x, int y, int z, double out[ORC_NDIR])
{
	size_t s[ORC_NDIR];
	uint32_t n = L->grid[lidx(L->dims, x, y, z)];
	list_node_slots(L, n, L->iteration % 2 == 0, s);
	for (int d = 0; d < ORC_NDIR; ++d) out[d] = L->pdf[s[d]];
}

void orc_list_set_node(orc_list_t * L, int x, int y, int z, const double in[ORC_NDIR])
{
	size_t s[ORC_NDIR];
	uint32_t n = L->grid[lidx(L->dims, x, y, z)];
	list_node_slots(L, n, L->iteration % 2 == 0, s);
	for (int d = 0; d < ORC_NDIR; ++d) L->pdf[s[d]] = in[d];
}

static void list_force(orc_list_t * L, double F)
{
	double ft[ORC_NDIR];
	for (int d = 0; d < ORC_NDIR; ++d) ft[d] = orc_force_term(d, F);
	const int X = L->dims[0], Y = L->dims[1], Z = L->dims[2];
	/* every slot belongs to exactly one node view: planes are independent */
	#pragma omp parallel for schedule(static)
	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		double v[ORC_NDIR];
		if (L->mask[lidx(L->dims, x, y, z)] == 0) continue;
		orc_list_get_node(L, x, y, z, v);
		for (int d = 0; d < ORC_NDIR; ++d) v[d] = v[d] + ft[d];
		orc_list_set_node(L, x, y, z, v);
	}
}

/* Per-node tree choice of the reference's vector split. */
static void list_tree_plan(const orc_list_t * L, int nThreads, int vsize,
                           unsigned char * evenTree, unsigned char * oddTree)
{
	const int nF = L->nFluid;
	memset(evenTree, ORC_TREE_SCALAR, nF);
	memset(oddTree, ORC_TREE_SCALAR, nF);
	if (vsize <= 1) return;
	for (int t = 0; t < nThreads; ++t) {            /* even: :233-238 */
		int s = thread_start(t, nF, nThreads), e = thread_start(t + 1, nF, nThreads);
		int nv = (e - s) - (e - s) % vsize;
		for (int n = s; n < s + nv; ++n) evenTree[n] = ORC_TREE_AVX;
	}
	/* odd: whole vectors of vsize inside each run, runs broken at thread
	 * boundaries (:578-754 with SetupConsecNodes :326-438) */
	int nRuns = orc_list_runs(L->adj, nF, nThreads, NULL);
	uint32_t * len = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nRuns);
	orc_list_runs(L->adj, nF, nThreads, len);
	int n = 0;
	for (int r = 0; r < nRuns; ++r) {
		int l = (int)len[r], nv = l - l % vsize;
		for (int i = 0; i < nv; ++i) oddTree[n + i] = ORC_TREE_AVX;
		n += l;
	}
	free(len);
}

void orc_list_run(orc_list_t * L, int iterations, double omega, double magic,
                  double F, int forceOn, int nThreads, int vsize)
{
	const double oe = omega, oo = orc_omega_odd(omega, magic);
	const size_t nc = (size_t)L->nCells;
	const int nF = L->nFluid;
	double * src = L->pdf;
	unsigned char * evenTree = (unsigned char *)malloc(nF > 0 ? nF : 1);
	unsigned char * oddTree  = (unsigned char *)malloc(nF > 0 ? nF : 1);
	list_tree_plan(L, nThreads > 0 ? nThreads : 1, vsize, evenTree, oddTree);

	for (int it = 0; it < iterations; it += 2) {
		/* even: local load, swapped local store (:380-465) */
		#pragma omp parallel for schedule(static)
		for (int n = 0; n < nF; ++n) {
			double p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d) p[d] = src[(size_t)d * nc + n];
			orc_trt_collide(p, q, oe, oo, evenTree[n]);
			for (int d = 0; d < ORC_NDIR; ++d) src[(size_t)ORC_INV[d] * nc + n] = q[d];
		}
		L->iteration = it;
		if (forceOn) list_force(L, F);

		/* odd: gather from inverse adjacency, scatter to adjacency (:755-845) */
		#pragma omp parallel for schedule(static)
		for (int n = 0; n < nF; ++n) {
			double p[ORC_NDIR], q[ORC_NDIR];
			const uint32_t * a = L->adj + (size_t)n * ORC_NLINK;
			for (int d = 0; d < ORC_NLINK; ++d) p[d] = src[a[ORC_INV[d]]];
			p[C_] = src[(size_t)C_ * nc + n];
			orc_trt_collide(p, q, oe, oo, oddTree[n]);
			for (int d = 0; d < ORC_NLINK; ++d) src[a[d]] = q[d];
			src[(size_t)C_ * nc + n] = q[C_];
		}
		L->iteration = it + 1;
		if (forceOn) list_force(L, F);
	}
	free(evenTree);
	free(oddTree);
}

void orc_list_dump(const orc_list_t * L, double * out)
{
	const int X = L->dims[0], Y = L->dims[1], Z = L->dims[2];
	size_t k = 0;
	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		if (L->mask[lidx(L->dims, x, y, z)] == 0) continue;
		orc_list_get_node(L, x, y, z, out + k * ORC_NDIR);
		++k;
	}
}

/* ------------------------------------------------------------------------
 * aa -- src/BenchKernelD3Q19AaCommon.c:309-637 (init, fix-up lists),
 *       src/BenchKernelD3Q19Aa.c:40-548 (steps),
 *       src/BenchKernelD3Q19AaCommon.c:127-215 (node view)
 * --------------------------------------------------------------------- */
static inline size_t p5(const orc_aa_t * A, int x, int y, int z, int d)
{
	/* ghost-layer coordinates, SoA, z fastest (src/BenchKernelD3Q19AaCommon.h:79-80) */
	const size_t gX = A->gdims[0], gY = A->gdims[1], gZ = A->gdims[2];
	return (size_t)d * gX * gY * gZ + (size_t)x * gY * gZ + (size_t)y * gZ + (size_t)z;
}

int orc_aa_init(orc_aa_t * A, const int * mask, const int dims[3], const int per[3])
{
	memset(A, 0, sizeof(*A));
	for (int k = 0; k < 3; ++k) { A->dims[k] = dims[k]; A->gdims[k] = dims[k] + 2; A->periodic[k] = per[k]; }
	A->mask = mask;
	A->iteration = -1;
	size_t n = (size_t)ORC_NDIR * A->gdims[0] * A->gdims[1] * A->gdims[2];
	A->pdf = (double *)calloc(n, sizeof(double));  /* interior 0.0 (:437-443); ghost cells read as 0.0 */
	if (!A->pdf) return -1;

	const int X = dims[0], Y = dims[1], Z = dims[2];
	/* one fix-up pair per (fluid cell, direction) whose push target is a
	 * solid cell or lies across a periodic face (:512-621) */
	int cap = 0;
	for (int pass = 0; pass < 2; ++pass) {
		int k = 0;
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (mask[lidx(dims, x, y, z)] == 0) continue;
			for (int d = 0; d < ORC_NDIR; ++d) {
				int nx = x + ORC_CX[d], ny = y + ORC_CY[d], nz = z + ORC_CZ[d];
				int oX = nx < 0 || nx >= X, oY = ny < 0 || ny >= Y, oZ = nz < 0 || nz >= Z;
				int src, dst;
				if ((oX && per[0]) || (oY && per[1]) || (oZ && per[2])) {
					int px = nx < 0 ? X - 1 : (nx >= X ? 0 : nx);
					int py = ny < 0 ? Y - 1 : (ny >= Y ? 0 : ny);
					int pz = nz < 0 ? Z - 1 : (nz >= Z ? 0 : nz);
					src = (int)p5(A, nx + 1, ny + 1, nz + 1, d);
					if (mask[lidx(dims, px, py, pz)] == 0) dst = (int)p5(A, x + 1, y + 1, z + 1, ORC_INV[d]);
					else                                   dst = (int)p5(A, px + 1, py + 1, pz + 1, d);
				}
				else if (oX || oY || oZ) continue;
				else if (mask[lidx(dims, nx, ny, nz)] == 0) {
					src = (int)p5(A, nx + 1, ny + 1, nz + 1, d);
					dst = (int)p5(A, x + 1, y + 1, z + 1, ORC_INV[d]);
				}
				else continue;
				if (pass == 1) { A->bbSrc[k] = src; A->bbDst[k] = dst; }
				++k;
			}
		}
		if (pass == 0) {
			cap = k;
			A->bbSrc = (int *)malloc(sizeof(int) * (size_t)(cap + 1));
			A->bbDst = (int *)malloc(sizeof(int) * (size_t)(cap + 1));
		}
	}
	A->nBB = cap;
	return 0;
}

void orc_aa_free(orc_aa_t * A)
{
	free(A->pdf); free(A->bbSrc); free(A->bbDst);
	memset(A, 0, sizeof(*A));
}

static void aa_node_slots(const orc_aa_t * A, int x, int y, int z, int parityEven, size_t s[ORC_NDIR])
{
	for (int d = 0; d < ORC_NDIR; ++d) {
		if (parityEven) s[d] = p5(A, x + 1 - ORC_CX[d], y + 1 - ORC_CY[d], z + 1 - ORC_CZ[d], ORC_INV[d]);
		else            s[d] = p5(A, x + 1, y + 1, z + 1, d);
	}
}

void orc_aa_get_node(const orc_aa_t * A, int x, int y, int z, double out[ORC_NDIR])
{
	size_t s[ORC_NDIR];
	aa_node_slots(A, x, y, z, A->iteration % 2 == 0, s);
	for (int d = 0; d < ORC_NDIR; ++d) out[d] = A->pdf[s[d]];
}

void orc_aa_set_node(orc_aa_t * A, int x, int y, int z, const double in[ORC_NDIR])
{
	size_t s[ORC_NDIR];
	aa_node_slots(A, x, y, z, A->iteration % 2 == 0, s);
	for (int d = 0; d < ORC_NDIR; ++d) A->pdf[s[d]] = in[d];
}

static void aa_force(orc_aa_t * A, double F)
{
	double ft[ORC_NDIR];
	for (int d = 0; d < ORC_NDIR; ++d) ft[d] = orc_force_term(d, F);
	#pragma omp parallel for schedule(static)
	for (int z = 0; z < A->dims[2]; ++z) for (int y = 0; y < A->dims[1]; ++y) for (int x = 0; x < A->dims[0]; ++x) {
		double v[ORC_NDIR];
		if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
		orc_aa_get_node(A, x, y, z, v);
		for (int d = 0; d < ORC_NDIR; ++d) v[d] = v[d] + ft[d];
		orc_aa_set_node(A, x, y, z, v);
	}
}

void orc_aa_run(orc_aa_t * A, int iterations, double omega, double magic, double F, int forceOn)
{
	const double oe = omega, oo = orc_omega_odd(omega, magic);
	const int X = A->dims[0], Y = A->dims[1], Z = A->dims[2];
	double * src = A->pdf;
	for (int it = 0; it < iterations; it += 2) {
		#pragma omp parallel for schedule(static)
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;      /* :182-184 */
			double p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d) p[d] = src[p5(A, x + 1, y + 1, z + 1, d)];
			orc_trt_collide(p, q, oe, oo, ORC_TREE_SCALAR);
			for (int d = 0; d < ORC_NDIR; ++d) src[p5(A, x + 1, y + 1, z + 1, ORC_INV[d])] = q[d];
		}
		for (int i = 0; i < A->nBB; ++i) src[A->bbSrc[i]] = src[A->bbDst[i]];   /* :309-311 */
		A->iteration = it;
		if (forceOn) aa_force(A, F);

		/* The reference also updates solid cells here (:380-491); what they
		 * write lands only in slots that the fix-up below overwrites or that
		 * no fluid cell reads, so fluid results do not depend on it. */
		#pragma omp parallel for schedule(static)
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
			double p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d)
				p[d] = src[p5(A, x + 1 - ORC_CX[d], y + 1 - ORC_CY[d], z + 1 - ORC_CZ[d], ORC_INV[d])];
			orc_trt_collide(p, q, oe, oo, ORC_TREE_SCALAR);
			for (int d = 0; d < ORC_NDIR; ++d)
				src[p5(A, x + 1 + ORC_CX[d], y + 1 + ORC_CY[d], z + 1 + ORC_CZ[d], d)] = q[d];
		}
		for (int i = 0; i < A->nBB; ++i) src[A->bbDst[i]] = src[A->bbSrc[i]];   /* :509-511 */
		A->iteration = it + 1;
		if (forceOn) aa_force(A, F);
	}
}

void orc_aa_dump(const orc_aa_t * A, double * out)
{
	size_t k = 0;
	for (int z = 0; z < A->dims[2]; ++z) for (int y = 0; y < A->dims[1]; ++y) for (int x = 0; x < A->dims[0]; ++x) {
		if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
		orc_aa_get_node(A, x, y, z, out + k * ORC_NDIR);
		++k;
	}
}

/* ------------------------------------------------------------------------
 * Initial state -- src/Kernel.c:258-438
 * --------------------------------------------------------------------- */
static void equilibrium_rest(double dens, double omega, double out[ORC_NDIR])
{
	const double w_0 = 1.0 / 3.0, w_1 = 1.0 / 18.0, w_2 = 1.0 / 36.0;
	const double ux = 0.0, uy = 0.0, uz = 0.0;
	const double ow0 = 3.0 * w_0 * omega, ow1 = 3.0 * w_1 * omega, ow2 = 3.0 * w_2 * omega;
	const double third = 1.0 / 3.0;
	const double t = third * dens - 0.5 * (ux * ux + uy * uy + uz * uz);
#define SQ(a) ((a) * (a))
	out[C_]  = ow0 * (t);
	out[NW_] = ow2 * (t - (ux - uy) + 1.5 * SQ(ux - uy));
	out[SE_] = ow2 * (t + (ux - uy) + 1.5 * SQ(ux - uy));
	out[NE_] = ow2 * (t + (ux + uy) + 1.5 * SQ(ux + uy));
	out[SW_] = ow2 * (t - (ux + uy) + 1.5 * SQ(ux + uy));
	out[TW_] = ow2 * (t - (ux - uz) + 1.5 * SQ(ux - uz));
	out[BE_] = ow2 * (t + (ux - uz) + 1.5 * SQ(ux - uz));
	out[TE_] = ow2 * (t + (ux + uz) + 1.5 * SQ(ux + uz));
	out[BW_] = ow2 * (t - (ux + uz) + 1.5 * SQ(ux + uz));
	out[TS_] = ow2 * (t - (uy - uz) + 1.5 * SQ(uy - uz));
	out[BN_] = ow2 * (t + (uy - uz) + 1.5 * SQ(uy - uz));
	out[TN_] = ow2 * (t + (uy + uz) + 1.5 * SQ(uy + uz));
	out[BS_] = ow2 * (t - (uy + uz) + 1.5 * SQ(uy + uz));
	out[N_]  = ow1 * (t + uy + 1.5 * SQ(uy));
	out[S_]  = ow1 * (t - uy + 1.5 * SQ(uy));
	out[E_]  = ow1 * (t + ux + 1.5 * SQ(ux));
	out[W_]  = ow1 * (t - ux + 1.5 * SQ(ux));
	out[T_]  = ow1 * (t + uz + 1.5 * SQ(uz));
	out[B_]  = ow1 * (t - uz + 1.5 * SQ(uz));
#undef SQ
}

void orc_initial_density_node(double rhoIn, double rhoOut, double omega, int x, int nX,
                              double out[ORC_NDIR])
{
	double dens = rhoIn + (rhoOut - rhoIn) * (x) / (nX - 1.0);
	equilibrium_rest(dens, omega, out);
}

void orc_initial_velocity_node(double omega, double out[ORC_NDIR])
{
	equilibrium_rest(1.0, omega, out);
}

/* ------------------------------------------------------------------------
 * Verification -- src/Kernel.c:456-625
 * --------------------------------------------------------------------- */
double orc_verify_norm(const double * ux, const int * fluid, int nY, int nZ,
                       double omega, double F)
{
	const double visc = (1.0 / omega - 0.5) / 3.0;
	int even = (nY % 2 == 0);
	int y = (nY - even - 1) / 2;
	double center = nY / 2.0;
	double dmin = (double)nY;
	double r2max = (dmin / 2.0 - 1.0) * (dmin / 2.0 - 1.0);
	double dev = 0.0;
	for (int z = 0; z < nZ; ++z) {
		double r2 = (z - center + 0.5) * (z - center + 0.5);
		double ana = (r2 >= r2max) ? 0.0 : F * (r2max - r2) / (2.0 * visc);
		double a0 = fluid[y * nZ + z] ? ux[y * nZ + z] + 0.5 * F : 0.0;
		double avg;
		if (even) {
			double a1 = fluid[(y + 1) * nZ + z] ? ux[(y + 1) * nZ + z] + 0.5 * F : 0.0;
			avg = (a0 + a1) / 2.0;
		} else avg = a0;
		if (ana != 0.0) {
			double r = (double)fabs(ana - avg) / ana;
			dev += r * r;
		}
	}
	return sqrt(dev);
}

/* ------------------------------------------------------------------------
 * Shifted storage, FP32 and scaled FP16.  No counterpart in the reference:
 * its single-precision build keeps f itself in float (src/Kernel.h:75-76).
 * Every slot holds the deviation s = f - w_d (w_d = w_inv(d), so a value
 * keeps its shift whichever direction reads it), the collision runs in
 * float on the deviations with the scalar tree of orc_trt_collide, and the
 * body-force increment is added in float when the next step loads a value,
 * or, after the last odd step of a run, when it is stored.  This schedule
 * is the definition of the variants: in FP32 it equals adding the increment
 * to the stored values after every step (same operands, same rounding); in
 * FP16 it avoids rounding the increment into binary16 every step.
 *
 * FP16 slots are modelled by their values: the float array holds
 * s * 2^k rounded to the nearest binary16 value, ties to even
 * (IEEE 754-2019 roundTiesToEven), and is read back as value * 2^-k.
 * --------------------------------------------------------------------- */
float orc_f16_round(float v)
{
	const double a = fabs((double)v);
	if (a == 0.0 || a != a || a > 1e300) return v;
	int e;
	(void)frexp(a, &e);                 /* a = m 2^e, m in [0.5, 1): exponent e - 1 */
	int ee = e - 1;
	if (ee < -14) ee = -14;             /* subnormal binary16: fixed quantum 2^-24 */
	const int q = ee - 10;              /* 10 fraction bits */
	const double r = ldexp(nearbyint(ldexp(a, -q)), q);   /* exact scaling, current rounding = to nearest even */
	if (r > 65504.0) return v < 0 ? -INFINITY : INFINITY;
	return (float)(v < 0 ? -r : r);
}

typedef struct {
	float oe, oo, w0, w1, w2, w1x3, w1nh, w2x3, w2nh;
} trt_f;

static trt_f trt_f_consts(double omega, double magic)
{
	/* each constant computed in double and rounded once */
	const double w1 = 1.0 / 18.0, w2 = 1.0 / 36.0;
	trt_f k = { (float)omega, (float)orc_omega_odd(omega, magic), (float)(1.0 / 3.0), (float)w1, (float)w2,
	            (float)(w1 * 3.0), (float)(w1 * 9.0 / 2.0), (float)(w2 * 3.0), (float)(w2 * 9.0 / 2.0) };
	return k;
}

/* orc_trt_collide's scalar tree in float (src/BenchKernelD3Q19ListAaPv.c:387-463) */
static void trt_collide_f(const float p[ORC_NDIR], float q[ORC_NDIR], const trt_f * k)
{
	const float ux = p[E_] + p[NE_] + p[SE_] + p[TE_] + p[BE_] - p[W_] - p[NW_] - p[SW_] - p[TW_] - p[BW_];
	const float uy = p[N_] + p[NE_] + p[NW_] + p[TN_] + p[BN_] - p[S_] - p[SE_] - p[SW_] - p[TS_] - p[BS_];
	const float uz = p[T_] + p[TE_] + p[TW_] + p[TN_] + p[TS_] - p[B_] - p[BE_] - p[BW_] - p[BN_] - p[BS_];
	const float dens = p[C_] + p[N_] + p[E_] + p[S_] + p[W_] + p[NE_] + p[SE_] + p[SW_] + p[NW_]
	                 + p[T_] + p[TN_] + p[TE_] + p[TS_] + p[TW_] + p[B_] + p[BN_] + p[BE_] + p[BS_] + p[BW_];
	const float dir = dens - (ux * ux + uy * uy + uz * uz) * 3.0f / 2.0f;

	q[C_] = p[C_] - k->oe * (p[C_] - k->w0 * dir);
	const float w1i = k->w1 * dir, w2i = k->w2 * dir;
	float ev, od, ui;
#define PAIR(a, b, uiexpr, wi, wnh, wx3)                               \
	ui = (uiexpr);                                                     \
	ev = k->oe * (0.5f * (p[a] + p[b]) - ui * ui * (wnh) - (wi));      \
	od = k->oo * (0.5f * (p[a] - p[b]) - ui * (wx3));                  \
	q[a] = p[a] - ev - od;                                             \
	q[b] = p[b] - ev + od;

	PAIR(N_,  S_,  uy,        w1i, k->w1nh, k->w1x3)
	PAIR(E_,  W_,  ux,        w1i, k->w1nh, k->w1x3)
	PAIR(T_,  B_,  uz,        w1i, k->w1nh, k->w1x3)
	PAIR(NW_, SE_, -ux + uy,  w2i, k->w2nh, k->w2x3)
	PAIR(NE_, SW_, ux + uy,   w2i, k->w2nh, k->w2x3)
	PAIR(TW_, BE_, -ux + uz,  w2i, k->w2nh, k->w2x3)
	PAIR(TE_, BW_, ux + uz,   w2i, k->w2nh, k->w2x3)
	PAIR(TS_, BN_, -uy + uz,  w2i, k->w2nh, k->w2x3)
	PAIR(TN_, BS_, uy + uz,   w2i, k->w2nh, k->w2x3)
#undef PAIR
}

typedef struct {
	int half;
	float scale, inv;
} codec_f;

static codec_f make_codec(int prec, int f16Shift)
{
	codec_f c = { prec == ORC_PREC_FP16S, 1.0f, 1.0f };
	if (c.half) { c.scale = ldexpf(1.0f, f16Shift); c.inv = ldexpf(1.0f, -f16Shift); }
	return c;
}
static inline float cload(const codec_f * c, float x) { return c->half ? x * c->inv : x; }
static inline float cstore(const codec_f * c, float v) { return c->half ? orc_f16_round(v * c->scale) : v; }

void orc_list_run_shifted(const orc_list_t * L, float * st, int prec, int f16Shift, int iterations,
                          double omega, double magic, double F, int forceOn)
{
	const trt_f k = trt_f_consts(omega, magic);
	const codec_f cd = make_codec(prec, f16Shift);
	const size_t nc = (size_t)L->nCells;
	const int nF = L->nFluid;
	float ft[ORC_NDIR];
	for (int d = 0; d < ORC_NDIR; ++d) ft[d] = (float)orc_force_term(d, F);

	for (int it = 0; it < iterations; it += 2) {
		const int loadForce = forceOn && it > 0, last = it + 2 >= iterations;
		#pragma omp parallel for schedule(static)
		for (int n = 0; n < nF; ++n) {
			float p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d) {
				p[d] = cload(&cd, st[(size_t)d * nc + n]);
				if (loadForce) p[d] = p[d] + ft[d];
			}
			trt_collide_f(p, q, &k);
			for (int d = 0; d < ORC_NDIR; ++d) st[(size_t)ORC_INV[d] * nc + n] = cstore(&cd, q[d]);
		}
		#pragma omp parallel for schedule(static)
		for (int n = 0; n < nF; ++n) {
			float p[ORC_NDIR], q[ORC_NDIR];
			const uint32_t * a = L->adj + (size_t)n * ORC_NLINK;
			for (int d = 0; d < ORC_NLINK; ++d) p[d] = cload(&cd, st[a[ORC_INV[d]]]);
			p[C_] = cload(&cd, st[(size_t)C_ * nc + n]);
			if (forceOn) for (int d = 0; d < ORC_NDIR; ++d) p[d] = p[d] + ft[d];
			trt_collide_f(p, q, &k);
			for (int d = 0; d < ORC_NLINK; ++d) {
				float v = q[d];
				if (forceOn && last) {
					/* a bounce-back store is read back by this node as inv(d) */
					const int bb = a[d] == (uint32_t)((size_t)ORC_INV[d] * nc + n);
					v = v + ft[bb ? ORC_INV[d] : d];
				}
				st[a[d]] = cstore(&cd, v);
			}
			st[(size_t)C_ * nc + n] = cstore(&cd, (forceOn && last) ? q[C_] + ft[C_] : q[C_]);
		}
	}
}

/* Node views of complete step pairs (local layout), in mask order. */
void orc_shifted_from_nodes(const orc_list_t * L, float * st, int prec, int f16Shift, const double * nodes)
{
	const codec_f cd = make_codec(prec, f16Shift);
	const int X = L->dims[0], Y = L->dims[1], Z = L->dims[2];
	size_t i = 0;
	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		if (L->mask[lidx(L->dims, x, y, z)] == 0) continue;
		const uint32_t n = L->grid[lidx(L->dims, x, y, z)];
		for (int d = 0; d < ORC_NDIR; ++d)
			st[(size_t)d * L->nCells + n] = cstore(&cd, (float)(nodes[i * ORC_NDIR + d] - ORC_W[d]));
		++i;
	}
}

void orc_shifted_to_nodes(const orc_list_t * L, const float * st, int prec, int f16Shift, double * nodes)
{
	const codec_f cd = make_codec(prec, f16Shift);
	const int X = L->dims[0], Y = L->dims[1], Z = L->dims[2];
	size_t i = 0;
	for (int z = 0; z < Z; ++z) for (int y = 0; y < Y; ++y) for (int x = 0; x < X; ++x) {
		if (L->mask[lidx(L->dims, x, y, z)] == 0) continue;
		const uint32_t n = L->grid[lidx(L->dims, x, y, z)];
		for (int d = 0; d < ORC_NDIR; ++d)
			nodes[i * ORC_NDIR + d] = ORC_W[d] + (double)cload(&cd, st[(size_t)d * L->nCells + n]);
		++i;
	}
}

/* aa-soa with shifted storage: the steps of orc_aa_run (even step, fix-up
 * copies, odd step, fix-up copies) over float slots laid out like A->pdf,
 * with the collision and the folded force increments of
 * orc_list_run_shifted.  The fix-ups move stored codes unchanged.  The last
 * odd step of a run keeps its collision outputs in float until the fix-ups
 * have put them in place, so that the node-view force increment is added
 * before the one rounding to the storage code; a node-view slot that the
 * step does not write (its writer would sit in the ghost layer) gets the
 * increment by decode, add, encode.  Slots outside node views that the last
 * step writes are stored without increment. */
int orc_aa_run_shifted(const orc_aa_t * A, float * st, int prec, int f16Shift, int iterations,
                       double omega, double magic, double F, int forceOn)
{
	const trt_f k = trt_f_consts(omega, magic);
	const codec_f cd = make_codec(prec, f16Shift);
	const int X = A->dims[0], Y = A->dims[1], Z = A->dims[2];
	const size_t n = (size_t)ORC_NDIR * A->gdims[0] * A->gdims[1] * A->gdims[2];
	float ft[ORC_NDIR];
	for (int d = 0; d < ORC_NDIR; ++d) ft[d] = (float)orc_force_term(d, F);
	float * raw = NULL;
	unsigned char * wr = NULL;

	for (int it = 0; it < iterations; it += 2) {
		const int loadForce = forceOn && it > 0, fold = forceOn && it + 2 >= iterations;
		#pragma omp parallel for schedule(static)
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
			float p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d) {
				p[d] = cload(&cd, st[p5(A, x + 1, y + 1, z + 1, d)]);
				if (loadForce) p[d] = p[d] + ft[d];
			}
			trt_collide_f(p, q, &k);
			for (int d = 0; d < ORC_NDIR; ++d) st[p5(A, x + 1, y + 1, z + 1, ORC_INV[d])] = cstore(&cd, q[d]);
		}
		for (int i = 0; i < A->nBB; ++i) st[A->bbSrc[i]] = st[A->bbDst[i]];

		if (fold && !raw) {
			raw = (float *)calloc(n, sizeof(float));
			wr = (unsigned char *)calloc(n, 1);
			if (!raw || !wr) { free(raw); free(wr); return -1; }
		}
		#pragma omp parallel for schedule(static)
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
			float p[ORC_NDIR], q[ORC_NDIR];
			for (int d = 0; d < ORC_NDIR; ++d) {
				p[d] = cload(&cd, st[p5(A, x + 1 - ORC_CX[d], y + 1 - ORC_CY[d], z + 1 - ORC_CZ[d], ORC_INV[d])]);
				if (forceOn) p[d] = p[d] + ft[d];
			}
			trt_collide_f(p, q, &k);
			for (int d = 0; d < ORC_NDIR; ++d) {
				const size_t s = p5(A, x + 1 + ORC_CX[d], y + 1 + ORC_CY[d], z + 1 + ORC_CZ[d], d);
				if (fold) { raw[s] = q[d]; wr[s] = 1; }
				else      st[s] = cstore(&cd, q[d]);
			}
		}
		if (!fold) {
			for (int i = 0; i < A->nBB; ++i) st[A->bbDst[i]] = st[A->bbSrc[i]];
			continue;
		}
		for (int i = 0; i < A->nBB; ++i) { raw[A->bbDst[i]] = raw[A->bbSrc[i]]; wr[A->bbDst[i]] = wr[A->bbSrc[i]]; }
		for (int x = 0; x < X; ++x) for (int y = 0; y < Y; ++y) for (int z = 0; z < Z; ++z) {
			if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
			for (int d = 0; d < ORC_NDIR; ++d) {
				const size_t s = p5(A, x + 1, y + 1, z + 1, d);
				st[s] = cstore(&cd, (wr[s] ? raw[s] : cload(&cd, st[s])) + ft[d]);
				wr[s] = 2;
			}
		}
		for (size_t s = 0; s < n; ++s) {
			if (wr[s] == 1) st[s] = cstore(&cd, raw[s]);
			wr[s] = 0;
		}
	}
	free(raw);
	free(wr);
	return 0;
}

/* Shifted storage of the aa layout: every slot at f = 0 (ghost cells
 * included), then the node views of the fluid cells (mask order) in their
 * local slots. */
void orc_aa_shifted_from_nodes(const orc_aa_t * A, float * st, int prec, int f16Shift, const double * nodes)
{
	const codec_f cd = make_codec(prec, f16Shift);
	const size_t plane = (size_t)A->gdims[0] * A->gdims[1] * A->gdims[2];
	for (int d = 0; d < ORC_NDIR; ++d) {
		const float zero = cstore(&cd, (float)(0.0 - ORC_W[d]));
		for (size_t i = 0; i < plane; ++i) st[(size_t)d * plane + i] = zero;
	}
	size_t i = 0;
	for (int z = 0; z < A->dims[2]; ++z) for (int y = 0; y < A->dims[1]; ++y) for (int x = 0; x < A->dims[0]; ++x) {
		if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
		for (int d = 0; d < ORC_NDIR; ++d)
			st[p5(A, x + 1, y + 1, z + 1, d)] = cstore(&cd, (float)(nodes[i * ORC_NDIR + d] - ORC_W[d]));
		++i;
	}
}

void orc_aa_shifted_to_nodes(const orc_aa_t * A, const float * st, int prec, int f16Shift, double * nodes)
{
	const codec_f cd = make_codec(prec, f16Shift);
	size_t i = 0;
	for (int z = 0; z < A->dims[2]; ++z) for (int y = 0; y < A->dims[1]; ++y) for (int x = 0; x < A->dims[0]; ++x) {
		if (A->mask[lidx(A->dims, x, y, z)] == 0) continue;
		for (int d = 0; d < ORC_NDIR; ++d)
			nodes[i * ORC_NDIR + d] = ORC_W[d] + (double)cload(&cd, st[p5(A, x + 1, y + 1, z + 1, d)]);
		++i;
	}
}
