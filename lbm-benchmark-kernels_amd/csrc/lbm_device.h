// lbm_device.h -- device-side D3Q19 TRT collision shared by the list and aa
// step kernels.
//
// The arithmetic is the reference's scalar path term by term
// (src/BenchKernelD3Q19ListAaPv.c:387-463, identical in
// src/BenchKernelD3Q19Aa.c:205-284): left-to-right sums for the moments,
// one rounding per operation.  The library is compiled with
// -ffp-contract=off so no FMA contraction changes the result; FP64 add/mul
// on CDNA are correctly rounded, which makes a step bit-identical to the
// reference's scalar evaluation of the same node.
#pragma once

#include <hip/hip_runtime.h>

#include <hip/hip_fp16.h>

#include "lbm_runtime.h"

namespace lbm {

// Direction slots (src/Kernel.h:89-111).
enum : int {
	dN = 0, dS, dE, dW, dNE, dSE, dNW, dSW, dT, dTN, dTE, dTW, dTS, dB, dBS, dBN, dBW, dBE, dC
};

__host__ __device__ constexpr int inv_dir(int d)
{
	return d == dN ? dS : d == dS ? dN : d == dE ? dW : d == dW ? dE :
	       d == dNE ? dSW : d == dSE ? dNW : d == dNW ? dSE : d == dSW ? dNE :
	       d == dT ? dB : d == dTN ? dBS : d == dTE ? dBW : d == dTW ? dBE :
	       d == dTS ? dBN : d == dB ? dT : d == dBS ? dTN : d == dBN ? dTS :
	       d == dBW ? dTE : d == dBE ? dTW : dC;
}
__host__ __device__ constexpr int cx(int d)
{
	return (d == dE || d == dNE || d == dSE || d == dTE || d == dBE) ? 1 :
	       (d == dW || d == dNW || d == dSW || d == dTW || d == dBW) ? -1 : 0;
}
__host__ __device__ constexpr int cy(int d)
{
	return (d == dN || d == dNE || d == dNW || d == dTN || d == dBN) ? 1 :
	       (d == dS || d == dSE || d == dSW || d == dTS || d == dBS) ? -1 : 0;
}
__host__ __device__ constexpr int cz(int d)
{
	return (d == dT || d == dTN || d == dTE || d == dTW || d == dTS) ? 1 :
	       (d == dB || d == dBS || d == dBN || d == dBW || d == dBE) ? -1 : 0;
}

// TRT constants in the compute precision R (values rounded from the
// double constants of TrtConsts).
template <class R>
struct TrtC {
	R oe, oo, w0, w1, w2, w1x3, w1nh, w2x3, w2nh;
};

template <class R>
inline TrtC<R> trt_cast(const TrtConsts & k)
{
	return TrtC<R>{(R)k.oe, (R)k.oo, (R)k.w0, (R)k.w1, (R)k.w2, (R)k.w1x3, (R)k.w1nh, (R)k.w2x3, (R)k.w2nh};
}

template <class R>
struct ForceT {
	R f[19];
};

template <class R>
inline ForceT<R> force_cast(const ForceTerms & t)
{
	ForceT<R> r;
	for (int d = 0; d < 19; ++d) r.f[d] = (R)t.f[d];
	return r;
}

// In place: f[] holds the 19 pre-collision values on entry and the
// post-collision value of each direction (not yet swapped) on exit.
//
// The same routine serves shifted storage: applied to the deviations
// s_d = f_d - w_d (rho0 = 1) it yields the post-collision deviations
// exactly, because the incompressible equilibrium is linear in the density
// (the velocity is sum c_d f_d without division by rho) and sum w_d = 1:
// the rho0 terms cancel in the rest, even and odd parts alike.
template <class R>
__device__ __forceinline__ void trt_collide(R (&f)[19], const TrtC<R> & k)
{
	const R ux = f[dE] + f[dNE] + f[dSE] + f[dTE] + f[dBE] - f[dW] - f[dNW] - f[dSW] - f[dTW] - f[dBW];
	const R uy = f[dN] + f[dNE] + f[dNW] + f[dTN] + f[dBN] - f[dS] - f[dSE] - f[dSW] - f[dTS] - f[dBS];
	const R uz = f[dT] + f[dTE] + f[dTW] + f[dTN] + f[dTS] - f[dB] - f[dBE] - f[dBW] - f[dBN] - f[dBS];
	const R dens = f[dC] + f[dN] + f[dE] + f[dS] + f[dW] + f[dNE] + f[dSE] + f[dSW] + f[dNW]
	             + f[dT] + f[dTN] + f[dTE] + f[dTS] + f[dTW] + f[dB] + f[dBN] + f[dBE] + f[dBS] + f[dBW];
	const R dir = dens - (ux * ux + uy * uy + uz * uz) * R(3.0) / R(2.0);

	f[dC] = f[dC] - k.oe * (f[dC] - k.w0 * dir);

	const R w1i = k.w1 * dir;
	const R w2i = k.w2 * dir;

	// One opposite pair (a, b) with projected velocity ui: the even part
	// relaxes with omegaEven, the odd part with omegaOdd.
	auto pair = [&](int a, int b, R ui, R wi, R wnh, R wx3) {
		const R ev = k.oe * (R(0.5) * (f[a] + f[b]) - ui * ui * wnh - wi);
		const R od = k.oo * (R(0.5) * (f[a] - f[b]) - ui * wx3);
		const R pa = f[a] - ev - od;
		const R pb = f[b] - ev + od;
		f[a] = pa;
		f[b] = pb;
	};
	pair(dN,  dS,  uy,       w1i, k.w1nh, k.w1x3);
	pair(dE,  dW,  ux,       w1i, k.w1nh, k.w1x3);
	pair(dT,  dB,  uz,       w1i, k.w1nh, k.w1x3);
	pair(dNW, dSE, -ux + uy, w2i, k.w2nh, k.w2x3);
	pair(dNE, dSW, ux + uy,  w2i, k.w2nh, k.w2x3);
	pair(dTW, dBE, -ux + uz, w2i, k.w2nh, k.w2x3);
	pair(dTE, dBW, ux + uz,  w2i, k.w2nh, k.w2x3);
	pair(dTS, dBN, -uy + uz, w2i, k.w2nh, k.w2x3);
	pair(dTN, dBS, uy + uz,  w2i, k.w2nh, k.w2x3);
}

}  // namespace lbm

namespace lbm {

// Storage codecs of the PDF array.  load/store convert between the stored
// value and the compute precision R; NT selects non-temporal accesses.  NT
// is a template argument so that each access is emitted with its hint (a
// run-time choice between two loads of the same address gets merged into
// one plain load).
template <bool NT, class T>
__device__ __forceinline__ T ld(const T * p)
{
	if constexpr (NT) return __builtin_nontemporal_load(p);
	else return *p;
}

template <bool NT, class T>
__device__ __forceinline__ void st(T * p, T v)
{
	if constexpr (NT) __builtin_nontemporal_store(v, p);
	else *p = v;
}

// dec/enc convert one stored value; load/store add the access.
struct CodecF64 {               // the reference's FP64 values
	using S = double;
	using R = double;
	static constexpr int VMAX = 1;   // nodes per 8-byte access
	__device__ __forceinline__ R dec(S v) const { return v; }
	__device__ __forceinline__ S enc(R v) const { return v; }
	template <bool NT> __device__ __forceinline__ R load(const S * p) const { return ld<NT>(p); }
	template <bool NT> __device__ __forceinline__ void store(S * p, R v) const { st<NT>(p, v); }
};

struct CodecF32 {               // deviation f - w in FP32
	using S = float;
	using R = float;
	static constexpr int VMAX = 2;
	__device__ __forceinline__ R dec(S v) const { return v; }
	__device__ __forceinline__ S enc(R v) const { return v; }
	template <bool NT> __device__ __forceinline__ R load(const S * p) const { return ld<NT>(p); }
	template <bool NT> __device__ __forceinline__ void store(S * p, R v) const { st<NT>(p, v); }
};

struct CodecF16 {               // deviation f - w times a power of two, in FP16
	using S = uint16_t;
	using R = float;
	static constexpr int VMAX = 4;
	float scale, inv;            // 2^k and 2^-k: the scaling itself is exact
	__device__ __forceinline__ R dec(S v) const { return __half2float(__ushort_as_half(v)) * inv; }
	__device__ __forceinline__ S enc(R v) const { return (uint16_t)__half_as_ushort(__float2half_rn(v * scale)); }
	template <bool NT> __device__ __forceinline__ R load(const S * p) const { return dec(ld<NT>(p)); }
	template <bool NT> __device__ __forceinline__ void store(S * p, R v) const { st<NT>(p, enc(v)); }
};

// V consecutive stored values as one access.
template <class S, int V>
using VecS = S __attribute__((ext_vector_type(V)));

}  // namespace lbm
