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

// In place: f[] holds the 19 pre-collision values on entry and the
// post-collision value of each direction (not yet swapped) on exit.
__device__ __forceinline__ void trt_collide(double (&f)[19], const TrtConsts & k)
{
	const double ux = f[dE] + f[dNE] + f[dSE] + f[dTE] + f[dBE] - f[dW] - f[dNW] - f[dSW] - f[dTW] - f[dBW];
	const double uy = f[dN] + f[dNE] + f[dNW] + f[dTN] + f[dBN] - f[dS] - f[dSE] - f[dSW] - f[dTS] - f[dBS];
	const double uz = f[dT] + f[dTE] + f[dTW] + f[dTN] + f[dTS] - f[dB] - f[dBE] - f[dBW] - f[dBN] - f[dBS];
	const double dens = f[dC] + f[dN] + f[dE] + f[dS] + f[dW] + f[dNE] + f[dSE] + f[dSW] + f[dNW]
	                  + f[dT] + f[dTN] + f[dTE] + f[dTS] + f[dTW] + f[dB] + f[dBN] + f[dBE] + f[dBS] + f[dBW];
	const double dir = dens - (ux * ux + uy * uy + uz * uz) * 3.0 / 2.0;

	f[dC] = f[dC] - k.oe * (f[dC] - k.w0 * dir);

	const double w1i = k.w1 * dir;
	const double w2i = k.w2 * dir;

	// One opposite pair (a, b) with projected velocity ui: the even part
	// relaxes with omegaEven, the odd part with omegaOdd.
	auto pair = [&](int a, int b, double ui, double wi, double wnh, double wx3) {
		const double ev = k.oe * (0.5 * (f[a] + f[b]) - ui * ui * wnh - wi);
		const double od = k.oo * (0.5 * (f[a] - f[b]) - ui * wx3);
		const double pa = f[a] - ev - od;
		const double pb = f[b] - ev + od;
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
