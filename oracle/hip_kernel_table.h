/*
 * hip_kernel_table.h -- registers the MI355X kernels of liblbmhip.so in the
 * reference driver's kernel table without editing any reference source.
 *
 * TEST INFRASTRUCTURE: used only by the "integration" target of
 * oracle/Makefile, which compiles the reference's own src/Main.c (unmodified)
 * with
 *
 *     gcc -include hip_kernel_table.h -I<reference>/src ... -c Main.c
 *
 * and links it with the reference's other objects and liblbmhip.so.  The
 * result is the reference's lbmbenchk with two more entries, selected like
 * every other kernel by `-kernel <name>` (src/Main.c:537-553):
 *
 *     list-aa-pv-soa-hip -> D3Q19ListAaPvInit_PushSoAHip / ...Deinit...
 *     aa-soa-hip         -> D3Q19AaInit_AaSoAHip / D3Q19AaDeinit_AaSoAHip
 *
 * How: this header is read before Main.c.  It includes the reference's
 * KernelFunctions.h with its table renamed, so the table (every entry of
 * src/KernelFunctions.h:56-188) and the header's include guard exist before
 * Main.c's own #include of that header, which then expands to nothing.  It
 * then defines the table Main.c iterates, g_kernels[], two entries longer,
 * and fills it before main() runs: the reference's entries in their order,
 * then the two above.  A maintainer who edits src/KernelFunctions.h instead
 * (the two declarations and two table entries shown in the integration
 * notes) gets the same table.
 */
#ifndef LBM_HIP_KERNEL_TABLE_H
#define LBM_HIP_KERNEL_TABLE_H

#include <string.h>

#define g_kernels lbm_reference_kernels
#include "KernelFunctions.h"
#undef g_kernels

/* include/lbm_hip.h declares the same four functions; it is not included
 * here because its ABI header restates the reference's types, which
 * Kernel.h has just defined. */
void D3Q19ListAaPvInit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params);
void D3Q19ListAaPvDeinit_PushSoAHip(LatticeDesc * ld, KernelData ** kernelData);
void D3Q19AaInit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData, Parameters * params);
void D3Q19AaDeinit_AaSoAHip(LatticeDesc * ld, KernelData ** kernelData);

#define LBM_REFERENCE_KERNELS (sizeof(lbm_reference_kernels) / sizeof(lbm_reference_kernels[0]))

static KernelFunctions g_kernels[LBM_REFERENCE_KERNELS + 2];

__attribute__((constructor)) static void lbm_hip_register_kernels(void)
{
	static const KernelFunctions hip[2] = {
		{ .Name = "list-aa-pv-soa-hip", .Init = D3Q19ListAaPvInit_PushSoAHip, .Deinit = D3Q19ListAaPvDeinit_PushSoAHip },
		{ .Name = "aa-soa-hip",         .Init = D3Q19AaInit_AaSoAHip,         .Deinit = D3Q19AaDeinit_AaSoAHip },
	};
	memcpy(g_kernels, lbm_reference_kernels, sizeof(lbm_reference_kernels));
	memcpy(g_kernels + LBM_REFERENCE_KERNELS, hip, sizeof(hip));
}

#endif
