/* Host helpers of liblbmhost.so (geometry generators, initial state, body
 * force, density, statistics, VTK output, verification) driven through a CPU
 * KernelData whose node views start at w_d; built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_sanitizers.py.  Writes its output
 * files into the working directory. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lbm_host.h"

static const double W[19] = {1./18,1./18,1./18,1./18,1./36,1./36,1./36,1./36,1./18,1./36,1./36,1./36,1./36,1./18,1./36,1./36,1./36,1./36,1./3};
static double * g_pdf; static int g_dims[3];
static size_t at(int x, int y, int z) { return ((size_t)z * g_dims[1] + y) * g_dims[0] + x; }
static void get_node(KernelData * kd, int x, int y, int z, PdfT * p) { (void)kd; memcpy(p, g_pdf + 19 * at(x, y, z), 19 * sizeof(double)); }
static void set_node(KernelData * kd, int x, int y, int z, PdfT * p) { (void)kd; memcpy(g_pdf + 19 * at(x, y, z), p, 19 * sizeof(double)); }

int main(void)
{
	const char * geos[] = {"box", "channel", "pipe", "blocks", "blocks-3", "fluid", "porous", "porous-3-0.7-5"};
	for (size_t gi = 0; gi < sizeof geos / sizeof geos[0]; ++gi) {
		int dims[3] = {40, 36, 34}, per[3] = {0, 0, 0};
		LatticeDesc ld;
		memset(&ld, 0, sizeof ld);
		GeoCreateByStr(geos[gi], dims, per, &ld);
		g_dims[0] = dims[0]; g_dims[1] = dims[1]; g_dims[2] = dims[2];
		g_pdf = (double *)calloc(19 * (size_t)dims[0] * dims[1] * dims[2], sizeof(double));
		for (size_t i = 0; i < (size_t)dims[0] * dims[1] * dims[2]; ++i) memcpy(g_pdf + 19 * i, W, sizeof W);
		KernelData kd;
		memset(&kd, 0, sizeof kd);
		kd.GetNode = get_node; kd.SetNode = set_node;
		kd.Dims[0] = dims[0]; kd.Dims[1] = dims[1]; kd.Dims[2] = dims[2];
		CaseData cd;
		memset(&cd, 0, sizeof cd);
		cd.Omega = 1.0; cd.XForce = 1e-5; cd.MaxIterations = 10; cd.VtkOutput = 1; cd.VtkModulus = 5; cd.StatisticsModulus = 1;
		KernelSetInitialDensity(&ld, &kd, &cd);
		KernelSetInitialVelocity(&ld, &kd, &cd);
		KernelAddBodyForce(&kd, &ld, &cd);
		printf("%s: nFluid %d density %f\n", geos[gi], ld.nFluid, (double)KernelDensity(&kd, &ld));
		KernelStatisticsAdv(&kd, &ld, &cd, 0, 1);
		KernelStatistics(&kd, &ld, &cd, 10);
		VtkWrite(&ld, &kd, &cd, 10);
		PdfT err = 0;
		KernelVerifiy(&ld, &kd, &cd, &err);
		printf("  verify norm %e\n", (double)err);
		free(g_pdf);
		GeoFree(&ld);
	}
	return 0;
}
