// lbm_runtime.hip -- host plumbing of liblbmhip.so (see lbm_runtime.h).
#include "lbm_runtime.h"

#include <errno.h>
#include <limits.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace lbm {

// Stencil of src/Kernel.h:121-140 (N S E W NE SE NW SW T TN TE TW TS B BS BN BW BE C).
const int kCx[19]  = { 0, 0, 1,-1, 1, 1,-1,-1, 0, 0, 1,-1, 0, 0, 0, 0,-1, 1, 0 };
const int kCy[19]  = { 1,-1, 0, 0, 1,-1, 1,-1, 0, 1, 0, 0,-1, 0,-1, 1, 0, 0, 0 };
const int kCz[19]  = { 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1,-1,-1,-1,-1,-1, 0 };
const int kInv[19] = { 1, 0, 3, 2, 7, 6, 5, 4,13,14,16,17,15, 8, 9,12,10,11,18 };

static int g_errorMode = LBM_HIP_ERRORS_EXIT;
static char g_lastError[1024] = "";

bool fail(const char * fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_lastError, sizeof(g_lastError), fmt, ap);
	va_end(ap);
	printf("ERROR: %s\n", g_lastError);
	fflush(stdout);
	if (g_errorMode == LBM_HIP_ERRORS_EXIT) exit(1);
	return false;
}

bool hip_ok(hipError_t e, const char * what)
{
	if (e == hipSuccess) return true;
	return fail("%s failed: %s (%s)", what, hipGetErrorName(e), hipGetErrorString(e));
}

// ---------------------------------------------------------------- padding
bool pad_parse(const char * s, PadSpec * p)
{
	*p = PadSpec();
	if (s == nullptr) return fail("padding string missing");
	if (!strncmp("auto", s, 4)) {
		// two entries: 16 MiB modulus + (2 MiB + 512 B), then 16 KiB + 128 B
		// (src/Padding.c:24-41)
		p->n = 2;
		p->mod[0] = 16 * 1024 * 1024; p->off[0] = 2 * 1024 * 1024 + 512;
		p->mod[1] = 16 * 1024;        p->off[1] = 128;
		return true;
	}
	if (!strncmp("no", s, 2) || !strncmp("off", s, 3)) return true;

	int commas = 0;
	for (const char * t = s; *t; ++t) commas += (*t == ',');
	if (commas + 1 > 16) return fail("too many entries in padding string.");
	int idx = 0, entries = 0;
	bool haveMod = false, haveOff = false;
	for (const char * t = s; *t;) {
		if (*t == ',') {
			if (!haveMod) return fail("modulus missing before next section in padding string.");
			haveMod = haveOff = false;
			++idx; ++t;
		} else if (*t == '+') {
			if (!haveMod) return fail("modulus missing before offset in padding string.");
			if (haveOff) return fail("offset is only allowed to be specified once per section in padding string.");
			++t;
		} else if (*t >= '0' && *t <= '9') {
			char * end = nullptr;
			errno = 0;
			long v = strtol(t, &end, 10);
			if (errno != 0 || v > INT_MAX) return fail("over- or underflow in modulus/offset of padding string.");
			if (!haveMod) { p->mod[idx] = (int)v; haveMod = true; ++entries; }
			else {
				if (haveOff) return fail("offset is only allowed to be specified once per section in padding string.");
				p->off[idx] = (int)v; haveOff = true;
			}
			t = end;
		} else {
			return fail("padding string contains invalid characters.");
		}
	}
	if (entries != commas + 1) return fail("did not find all padding entries.");
	for (int i = 0; i < entries; ++i)
		if (p->off[i] >= p->mod[i]) return fail("offset in padding entry %d is equal or larger than modulus.", i);
	p->n = entries;
	return true;
}

long pad_cells(long n, int cellBytes, const PadSpec * ps)
{
	PadSpec autoSpec;
	if (ps == nullptr) { pad_parse("auto", &autoSpec); ps = &autoSpec; }
	for (int i = 0; i < ps->n; ++i) {
		long add;
		if (ps->mod[i] == 0) {
			add = (long)ps->off[i] * cellBytes;    // the reference scales the bare offset by the cell size
		} else {
			long m = ps->mod[i] / cellBytes;
			if (m == 0) {
				fail("modulus of %d byte in padding entry %d becomes zero for PDF size %d byte.",
				     ps->mod[i], i, cellBytes);
				return -1;
			}
			long o = ps->off[i] / cellBytes;
			add = (o + m - n % m) % m;
		}
		n += add;
	}
	return n;
}

// ------------------------------------------------------------ kernel args
static bool arg_int(const Parameters * p, int & i, int * out, const char * name, int lo, int hi)
{
	if (i + 1 >= p->nKernelArgs) return fail("argument %s requires a parameter.", name);
	char * end = nullptr;
	long v = strtol(p->KernelArgs[++i], &end, 0);
	if (end == p->KernelArgs[i] || *end != '\0') return fail("argument %s expects an integer.", name);
	if (v < lo || v > hi) return fail("argument %s must be in [%d, %d].", name, lo, hi);
	*out = (int)v;
	return true;
}

bool parse_kernel_args(const Parameters * p, KernelArgs * a, bool listKernel)
{
	*a = KernelArgs();
	if (p == nullptr) return true;
	for (int i = 0; i < p->nKernelArgs; ++i) {
		const char * s = p->KernelArgs[i];
		auto is = [&](const char * x) { return !strcmp(s, x); };
		if (is("-blk") || is("--blk")) {
			int v;
			if (!arg_int(p, i, &v, "-blk", 0, INT_MAX)) return false;
			a->blk[0] = a->blk[1] = a->blk[2] = v;
		} else if (is("-blk-x") || is("--blk-x")) {
			if (!arg_int(p, i, &a->blk[0], "-blk-x", 0, INT_MAX)) return false;
		} else if (is("-blk-y") || is("--blk-y")) {
			if (!arg_int(p, i, &a->blk[1], "-blk-y", 0, INT_MAX)) return false;
		} else if (is("-blk-z") || is("--blk-z")) {
			if (!arg_int(p, i, &a->blk[2], "-blk-z", 0, INT_MAX)) return false;
		} else if (listKernel && (is("-pad") || is("--pad"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument %s requires a parameter.", s);
			if (!pad_parse(p->KernelArgs[++i], &a->pad)) return false;
			a->havePad = true;
		} else if (is("-force") || is("--force")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -force requires on or off.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "on")) a->forceOn = 1;
			else if (!strcmp(v, "off")) a->forceOn = 0;
			else return fail("argument -force requires on or off.");
		} else if (is("-lambda") || is("--lambda")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -lambda requires a value.");
			char * end = nullptr;
			double v = strtod(p->KernelArgs[++i], &end);
			if (end == p->KernelArgs[i] || *end != '\0' || !(v > 0.0)) return fail("argument -lambda expects a positive number.");
			a->magic = v;
		} else if (listKernel && (is("-adj") || is("--adj"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -adj requires rle or full.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "rle")) a->adjRle = 1;
			else if (!strcmp(v, "full")) a->adjRle = 0;
			else return fail("argument -adj requires rle or full.");
		} else if (is("-nt") || is("--nt")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -nt requires on or off.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "on")) a->nontemporal = 1;
			else if (!strcmp(v, "off")) a->nontemporal = 0;
			else return fail("argument -nt requires on or off.");
		} else if (is("-prec") || is("--prec")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -prec requires fp64, fp32s or fp16s.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "fp64")) a->prec = PREC_FP64;
			else if (!strcmp(v, "fp32s")) a->prec = PREC_FP32S;
			else if (!strcmp(v, "fp16s")) a->prec = PREC_FP16S;
			else return fail("argument -prec requires fp64, fp32s or fp16s.");
		} else if (is("-fp16-scale") || is("--fp16-scale")) {
			int v;
			if (!arg_int(p, i, &v, "-fp16-scale", 1, 1 << 24)) return false;
			if (v & (v - 1)) return fail("argument -fp16-scale must be a power of two.");
			int k = 0;
			while ((1 << k) < v) ++k;
			a->f16Shift = k;
		} else if (is("-vec") || is("--vec")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -vec requires auto or 1.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "auto")) a->evenVec = 1;
			else if (!strcmp(v, "1")) a->evenVec = 0;
			else return fail("argument -vec requires auto or 1.");
		} else if (!listKernel && (is("-odd-pairs") || is("--odd-pairs"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -odd-pairs requires auto, on or off.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "auto")) a->oddPairs = -1;
			else if (!strcmp(v, "on")) a->oddPairs = 1;
			else if (!strcmp(v, "off")) a->oddPairs = 0;
			else return fail("argument -odd-pairs requires auto, on or off.");
		} else if (listKernel && (is("-shadow") || is("--shadow"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -shadow requires eager or lazy.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "eager")) a->lazyShadow = 0;
			else if (!strcmp(v, "lazy")) a->lazyShadow = 1;
			else return fail("argument -shadow requires eager or lazy.");
		} else if (listKernel && (is("-odd-order") || is("--odd-order"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -odd-order requires band, group or plain.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "plain")) a->oddOrder = 0;
			else if (!strcmp(v, "group")) a->oddOrder = 1;
			else if (!strcmp(v, "band")) a->oddOrder = 2;
			else return fail("argument -odd-order requires band, group or plain.");
		} else if (listKernel && (is("-even-order") || is("--even-order"))) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -even-order requires down or up.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "down")) a->evenDown = 1;
			else if (!strcmp(v, "up")) a->evenDown = 0;
			else return fail("argument -even-order requires down or up.");
		} else if (is("-statistics") || is("--statistics")) {
			a->statistics = 1;
		} else if (is("-launch-events") || is("--launch-events")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -launch-events requires on or off.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "on")) a->launchEvents = 1;
			else if (!strcmp(v, "off")) a->launchEvents = 0;
			else return fail("argument -launch-events requires on or off.");
		} else if (is("-graph") || is("--graph")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -graph requires on or off.");
			const char * v = p->KernelArgs[++i];
			if (!strcmp(v, "on")) a->graph = 1;
			else if (!strcmp(v, "off")) a->graph = 0;
			else return fail("argument -graph requires on or off.");
		} else if (is("-quiet") || is("--quiet")) {
			a->quiet = 1;
		} else if (is("-even-cap") || is("--even-cap")) {
			if (i + 1 >= p->nKernelArgs) return fail("argument -even-cap requires on, off or force.");
			const char * v = p->KernelArgs[++i];
			a->evenCapGiven = 1;
			if (!strcmp(v, "on")) a->evenCap = 1;
			else if (!strcmp(v, "off")) a->evenCap = 0;
			else if (!strcmp(v, "force")) a->evenCap = 2;
			else return fail("argument -even-cap requires on, off or force.");
		} else if (is("-bs") || is("--bs")) {
			if (!arg_int(p, i, &a->blockSize, "-bs", 64, 512)) return false;
			if (a->blockSize % 64) return fail("argument -bs must be a multiple of 64.");
		} else if (is("-h") || is("-help") || is("--help")) {
			printf("Kernel parameters:\n");
			printf("  [-blk <n>] [-blk-[xyz] <n>]\n");
			if (listKernel) printf("  [-pad auto|modulus_1+offset_1(,modulus_n+offset_n)*]\n");
			printf("  [-prec fp64|fp32s|fp16s] [-fp16-scale <2^k>] [-vec auto|1] [-statistics] [-graph on|off]\n");
			if (listKernel) printf("  [-adj rle|full] [-shadow eager|lazy] [-odd-order band|group|plain] [-even-order down|up]\n");
			else printf("  [-odd-pairs auto|on|off]\n");
			printf("  [-force on|off] [-lambda <magic>] [-bs <threads per block>] [-nt on|off] [-even-cap on|off|force] [-quiet]\n");
			printf("  [-launch-events on|off]\n");
			return fail("kernel help requested.");
		} else {
			return fail("unknown kernel parameter: %s.", s);
		}
	}
	return true;
}

// ---------------------------------------------------------------- binary16
uint16_t f32_to_f16(float f)
{
	uint32_t x;
	memcpy(&x, &f, 4);
	const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
	if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
	if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          // rounds past 65504
	if (ax >= 0x38800000u) {                                             // normal result
		uint32_t h = ((((ax >> 23) - 127u + 15u)) << 10) | ((ax & 0x7fffffu) >> 13);
		const uint32_t rem = ax & 0x1fffu;
		if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;       // carries into the exponent
		return (uint16_t)(sign | h);
	}
	if (ax <= 0x33000000u) return (uint16_t)sign;                       // <= 2^-25: ties to zero
	const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
	const uint32_t shift = 126u - e;                                     // subnormal: m * 2^(e-150) / 2^-24
	uint32_t h = m >> shift;
	const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
	if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
	return (uint16_t)(sign | h);
}

float f16_to_f32(uint16_t h)
{
	const uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
	uint32_t x;
	if (e == 0x1fu) x = sign | 0x7f800000u | (m << 13);
	else if (e != 0) x = sign | ((e + 112u) << 23) | (m << 13);
	else if (m == 0) x = sign;
	else {
		int k = 0;
		uint32_t mm = m;
		while (!(mm & 0x400u)) { mm <<= 1; ++k; }
		x = sign | ((uint32_t)(113 - k) << 23) | ((mm & 0x3ffu) << 13);
	}
	float f;
	memcpy(&f, &x, 4);
	return f;
}

// ---------------------------------------------------------------- physics
TrtConsts trt_consts(double omega, double magic)
{
	TrtConsts k;
	k.oe = omega;
	k.oo = 1.0 / (0.5 + magic / (1.0 / omega - 0.5));
	k.w0 = 1.0 / 3.0;
	k.w1 = 1.0 / 18.0;
	k.w2 = 1.0 / 36.0;
	k.w1x3 = k.w1 * 3.0;
	k.w1nh = k.w1 * 9.0 / 2.0;
	k.w2x3 = k.w2 * 3.0;
	k.w2nh = k.w2 * 9.0 / 2.0;
	return k;
}

ForceTerms force_terms(double F)
{
	static const double w[19] = {
		1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
		1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
		1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 3.0 };
	ForceTerms t;
	for (int d = 0; d < 19; ++d) t.f[d] = 3.0 * w[d] * kCx[d] * F;
	return t;
}

// ----------------------------------------------------------- common state
bool shadow_writable(const Common * c)
{
	if (!c->shadowStale) return true;
	return fail("SetNode: the device lattice is newer than the host copy (shadow download off); "
	            "call LbmHipSyncShadow first.");
}

bool common_init(Common * c, const Parameters * params, bool listKernel)
{
	if (!parse_kernel_args(params, &c->args, listKernel)) return false;
	c->launchEvents = c->args.launchEvents;
	int n = 0;
	const hipError_t e = hipGetDeviceCount(&n);
	if (e == hipErrorNoDevice || (e == hipSuccess && n < 1))
		return fail("no HIP device found; the list/aa kernels of this library run on the GPU only.");
	if (!hip_ok(e, "hipGetDeviceCount")) return false;
	if (!hip_ok(hipGetDevice(&c->device), "hipGetDevice")) return false;
	if (!hip_ok(hipDeviceGetAttribute(&c->ldsPerCU, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device),
	            "hipDeviceGetAttribute")) return false;
	if (!hip_ok(hipDeviceGetAttribute(&c->ldsPerBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device),
	            "hipDeviceGetAttribute")) return false;
	if (!hip_ok(hipDeviceGetAttribute(&c->nCU, hipDeviceAttributeMultiprocessorCount, c->device),
	            "hipDeviceGetAttribute")) return false;
	if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
	if (!hip_ok(hipEventCreate(&c->evLoop[0]), "hipEventCreate")) return false;
	if (!hip_ok(hipEventCreate(&c->evLoop[1]), "hipEventCreate")) return false;
	return true;
}

void common_graph_drop(Common * c)
{
	if (c->graphExec) (void)hipGraphExecDestroy(c->graphExec);
	if (c->graph) (void)hipGraphDestroy(c->graph);
	c->graphExec = nullptr;
	c->graph = nullptr;
	c->graphSteps = 0;
}

void common_free(Common * c)
{
	common_graph_drop(c);
	for (int i = 0; i < c->nEv; ++i) (void)hipEventDestroy(c->ev[i]);
	free(c->ev);
	c->ev = nullptr;
	c->nEv = 0;
	for (auto & e : c->evLoop) if (e) { (void)hipEventDestroy(e); e = nullptr; }
	if (c->stream) { (void)hipStreamDestroy(c->stream); c->stream = nullptr; }
}

StreamShape stream_shape(const Common * c, int defaultBlock, uint64_t threads)
{
	StreamShape sh = {defaultBlock, 0};
	if (!c->args.evenCap) return sh;
	if (c->args.evenCap != 2 && threads < 16ull * 512 * (uint64_t)(c->nCU > 0 ? c->nCU : 1)) return sh;
	const size_t want = (size_t)c->ldsPerCU / 2 + 8192;
	if (c->ldsPerCU <= 0 || want > (size_t)c->ldsPerBlock) return sh;
	sh.block = 512;
	sh.lds = want;
	return sh;
}

bool common_events(Common * c, int launches)
{
	int need = 2 * launches;
	if (need <= c->nEv) return true;
	hipEvent_t * ev = (hipEvent_t *)realloc(c->ev, sizeof(hipEvent_t) * (size_t)need);
	if (!ev) return fail("allocating %d events failed.", need);
	c->ev = ev;
	for (int i = c->nEv; i < need; ++i) {
		// timing only: no system-scope fence (its cache writeback and
		// invalidation would otherwise slow the launch that follows; the
		// host reads results only after a stream synchronisation)
		if (!hip_ok(hipEventCreateWithFlags(&c->ev[i], hipEventDisableSystemFence), "hipEventCreateWithFlags")) {
			c->nEv = i;
			return false;
		}
	}
	c->nEv = need;
	return true;
}

void common_collect(Common * c, int launches, bool perLaunch)
{
	float ms = 0.0f;
	(void)hipEventElapsedTime(&ms, c->evLoop[0], c->evLoop[1]);
	c->t.loopMs = ms;
	c->t.steps = launches;
	c->t.evenMsAvg = c->t.oddMsAvg = 0.0;
	c->t.evenMsMin = c->t.oddMsMin = 0.0;
	if (!perLaunch || launches == 0) { (void)hipGetLastError(); return; }
	double se = 0, so = 0, me = 1e30, mo = 1e30;
	int ne = 0, no = 0;
	for (int i = 0; i < launches; ++i) {
		float m = 0.0f;
		(void)hipEventElapsedTime(&m, c->ev[2 * i], c->ev[2 * i + 1]);
		if (i % 2 == 0) { se += m; ++ne; if (m < me) me = m; }
		else            { so += m; ++no; if (m < mo) mo = m; }
	}
	c->t.evenMsAvg = ne ? se / ne : 0.0;
	c->t.oddMsAvg = no ? so / no : 0.0;
	c->t.evenMsMin = ne ? me : 0.0;
	c->t.oddMsMin = no ? mo : 0.0;
	(void)hipGetLastError();
}

void * host_alloc(size_t bytes)
{
	void * p = nullptr;
	if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
	return p;
}

void host_free(void * p)
{
	if (p) (void)hipHostFree(p);
}

}  // namespace lbm

// -------------------------------------------------------------- C exports
extern "C" {

void LbmHipSetErrorMode(int mode)
{
	lbm::g_errorMode = (mode == LBM_HIP_ERRORS_RETURN) ? LBM_HIP_ERRORS_RETURN : LBM_HIP_ERRORS_EXIT;
}

const char * LbmHipLastError(void) { return lbm::g_lastError; }

void LbmHipClearError(void) { lbm::g_lastError[0] = '\0'; }

int LbmHipDeviceInfo(char * name, int nameLen, int * cus, double * memClockGHz)
{
	// a query: reports -1 without the error path when no device is present
	int dev = 0, n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return -1;
	if (hipGetDevice(&dev) != hipSuccess) return -1;
	hipDeviceProp_t p;
	if (hipGetDeviceProperties(&p, dev) != hipSuccess) return -1;
	if (name && nameLen > 0) {
		if (p.name[0]) snprintf(name, (size_t)nameLen, "%s (%s)", p.name, p.gcnArchName);
		else snprintf(name, (size_t)nameLen, "%s, %d CUs", p.gcnArchName, p.multiProcessorCount);
	}
	if (cus) *cus = p.multiProcessorCount;
	if (memClockGHz) *memClockGHz = p.memoryClockRate / 1.0e6;
	return 0;
}

}  // extern "C"

// ----------------------------------------------------------- kernel table
// Laid out like the reference's g_kernels[] (src/KernelFunctions.h:56-188);
// a driver looks kernels up by name with strcasecmp (src/Main.c:537-553).
extern "C" {
KernelFunctions g_lbmHipKernels[] = {
	{ "list-aa-pv-soa", D3Q19ListAaPvInit_PushSoAHip, D3Q19ListAaPvDeinit_PushSoAHip },
	{ "aa-soa",         D3Q19AaInit_AaSoAHip,         D3Q19AaDeinit_AaSoAHip },
};
const int g_lbmHipKernelCount = (int)(sizeof(g_lbmHipKernels) / sizeof(g_lbmHipKernels[0]));
}

namespace lbm {
Common * common_of(KernelData * kd)
{
	if (!kd) return nullptr;
	Common * c = list_common(kd);
	return c ? c : aa_common(kd);
}
}  // namespace lbm

extern "C" {

int LbmHipGetTimings(const KernelData * kd, LbmHipTimings * out)
{
	lbm::Common * c = lbm::common_of(const_cast<KernelData *>(kd));
	if (!c || !out) { lbm::fail("LbmHipGetTimings: not a kernel of this library."); return -1; }
	*out = c->t;
	return 0;
}

int LbmHipSetLaunchEvents(KernelData * kd, int on)
{
	lbm::Common * c = lbm::common_of(kd);
	if (!c) { lbm::fail("LbmHipSetLaunchEvents: not a kernel of this library."); return -1; }
	c->launchEvents = on != 0;
	return 0;
}

int LbmHipSetShadowDownload(KernelData * kd, int on)
{
	lbm::Common * c = lbm::common_of(kd);
	if (!c) { lbm::fail("LbmHipSetShadowDownload: not a kernel of this library."); return -1; }
	c->shadowDownload = on != 0;
	return 0;
}

int LbmHipSyncShadow(KernelData * kd)
{
	lbm::Common * c = lbm::common_of(kd);
	if (!c || !c->download) { lbm::fail("LbmHipSyncShadow: not a kernel of this library."); return -1; }
	return c->download(kd) ? 0 : -1;
}

}  // extern "C"

extern "C" {

int LbmHipDumpNodes(KernelData * kd, const LatticeDesc * ld, PdfT * out)
{
	if (!lbm::common_of(kd) || !ld || !out) { lbm::fail("LbmHipDumpNodes: invalid arguments."); return -1; }
	size_t k = 0;
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < ld->Dims[0]; ++x) {
		if (ld->Lattice[((size_t)z * ld->Dims[1] + y) * ld->Dims[0] + x] == LBM_CELL_OBSTACLE) continue;
		kd->GetNode(kd, x, y, z, out + 19 * k);
		++k;
	}
	return 0;
}

int LbmHipLoadNodes(KernelData * kd, const LatticeDesc * ld, const PdfT * in)
{
	if (!lbm::common_of(kd) || !ld || !in) { lbm::fail("LbmHipLoadNodes: invalid arguments."); return -1; }
	size_t k = 0;
	for (int z = 0; z < ld->Dims[2]; ++z)
	for (int y = 0; y < ld->Dims[1]; ++y)
	for (int x = 0; x < ld->Dims[0]; ++x) {
		if (ld->Lattice[((size_t)z * ld->Dims[1] + y) * ld->Dims[0] + x] == LBM_CELL_OBSTACLE) continue;
		kd->SetNode(kd, x, y, z, const_cast<PdfT *>(in + 19 * k));
		++k;
	}
	return 0;
}

}  // extern "C"

// ------------------------------------------------------- stream yardsticks
namespace lbm {

__global__ __launch_bounds__(256) void k_stream_copy(double * __restrict__ dst, const double * __restrict__ src, uint64_t n)
{
	const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n) dst[i] = src[i];
}

// 19 planes read and written in place, the memory pattern of an even step
__global__ __launch_bounds__(256) void k_stream_update19(double * p, uint32_t n)
{
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	double v[19];
#pragma unroll
	for (int d = 0; d < 19; ++d) v[d] = p[(size_t)d * n + i];
#pragma unroll
	for (int d = 0; d < 19; ++d) p[(size_t)d * n + i] = v[d] * 0.5 + 1.0;
}

}  // namespace lbm

extern "C" int LbmHipStreamBench(uint64_t nDoubles, uint32_t nNodes, int reps, LbmHipStreamResult * out)
{
	using namespace lbm;
	if (!out || nDoubles == 0 || nNodes == 0 || reps < 1) { fail("LbmHipStreamBench: invalid arguments."); return -1; }
	double * a = nullptr;
	double * b = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int rc = -1;
	do {
		const size_t bytesA = sizeof(double) * (size_t)(nDoubles > 19ull * nNodes ? nDoubles : 19ull * nNodes);
		if (!hip_ok(hipMalloc((void **)&a, bytesA), "hipMalloc")) break;
		if (!hip_ok(hipMalloc((void **)&b, sizeof(double) * (size_t)nDoubles), "hipMalloc")) break;
		if (!hip_ok(hipMemset(a, 0, bytesA), "hipMemset")) break;
		if (!hip_ok(hipEventCreate(&e0), "hipEventCreate") || !hip_ok(hipEventCreate(&e1), "hipEventCreate")) break;
		const unsigned gc = (unsigned)((nDoubles + 255) / 256), gu = (nNodes + 255) / 256;
		float ms = 0.0f;
		(void)hipGetLastError();
		for (int i = 0; i < 2; ++i) k_stream_copy<<<gc, 256>>>(b, a, nDoubles);
		if (!hip_ok(hipEventRecord(e0), "hipEventRecord")) break;
		for (int i = 0; i < reps; ++i) k_stream_copy<<<gc, 256>>>(b, a, nDoubles);
		if (!hip_ok(hipEventRecord(e1), "hipEventRecord") || !hip_ok(hipEventSynchronize(e1), "copy kernel")) break;
		if (!hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) break;
		out->copyBytes = 16.0 * (double)nDoubles;
		out->copyGBs = out->copyBytes * reps / (ms * 1e-3) / 1e9;
		for (int i = 0; i < 2; ++i) k_stream_update19<<<gu, 256>>>(a, nNodes);
		if (!hip_ok(hipEventRecord(e0), "hipEventRecord")) break;
		for (int i = 0; i < reps; ++i) k_stream_update19<<<gu, 256>>>(a, nNodes);
		if (!hip_ok(hipEventRecord(e1), "hipEventRecord") || !hip_ok(hipEventSynchronize(e1), "update-19 kernel")) break;
		if (!hip_ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) break;
		if (!hip_ok(hipGetLastError(), "stream kernels")) break;
		out->update19Bytes = 2.0 * 19 * 8 * (double)nNodes;
		out->update19GBs = out->update19Bytes * reps / (ms * 1e-3) / 1e9;
		rc = 0;
	} while (0);
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipFree(a);
	(void)hipFree(b);
	return rc;
}
