/*
 * main.c -- lbmbenchk-hip, the benchmark driver for the GPU kernels.
 *
 * Same command line, run sequence and output lines as the reference driver
 * (src/Main.c:127-652): parse flags, build the geometry, look the kernel up
 * by name, Init -> [initial state for -verify] -> Kernel -> statistics ->
 * [verification] -> metrics -> Deinit, print the "P:" line, exit 1 when the
 * verification norm exceeds 0.1.  Kernels come from liblbmhip.so's table;
 * the helpers from liblbmhost.so.
 */
#include "lbm_hip.h"
#include "lbm_host.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

static int parse_dims(const char * s, int d[3])
{
	/* "XxYxZ" or a single number for a cube */
	int n = sscanf(s, "%dx%dx%d", &d[0], &d[1], &d[2]);
	if (n == 1) { d[1] = d[2] = d[0]; n = 3; }
	if (n != 3 || d[0] <= 0 || d[1] <= 0 || d[2] <= 0) {
		printf("ERROR: invalid dimensions \"%s\", expected XxYxZ.\n", s);
		return 0;
	}
	return 1;
}

static void usage(void)
{
	printf("Usage:\n");
	printf("./lbmbenchk-hip -list\n");
	printf("./lbmbenchk-hip\n");
	printf("      [-dims XxYxZ] [-geometry box|channel|pipe|blocks[-<n>]|fluid|porous[-<r>[-<porosity>[-<seed>]]]]\n");
	printf("      [-iterations <iterations>] [-rho-in <density>] [-rho-out <density>] [-omega <omega>]\n");
	printf("      [-x-force <force>] [-kernel <kernel>] [-periodic-x] [-periodic-y] [-periodic-z]\n");
	printf("      [-t <host threads>] [-verify] [-vtk [<every-n-iteration>]] [-statistics <every-n-iteration>]\n");
	printf("      -- <kernel specific parameters>\n");
}

int main(int argc, char * argv[])
{
	int dims[3] = { 20, 20, 20 };
	const char * geometry = "channel";
	const char * kernelName = "list-aa-pv-soa";
	int verify = 0;
	int nThreads = 1;
	int statistics = 0;
	int periodic[3] = { 0, 0, 0 };
	CaseData cd;
	memset(&cd, 0, sizeof(cd));
	cd.MaxIterations = 10;
	cd.RhoIn = 1.0;
	cd.RhoOut = 1.0;
	cd.Omega = 1.0;
	cd.VtkModulus = 100;
	cd.StatisticsModulus = 100;
	cd.XForce = 0.00001;

	Parameters p;
	p.nArgs = argc;
	p.Args = argv;
	p.nKernelArgs = 0;
	p.KernelArgs = NULL;

	printf("# LBM Benchmark Kernels, MI355X build (HIP, gfx950), type: verification\n");

	for (int i = 1; i < argc; ++i) {
		const char * a = argv[i];
#define IS(x) (!strcmp(a, x) || !strcmp(a, "-" x))
#define NEED() do { if (i + 1 >= argc) { printf("ERROR: argument %s requires a parameter.\n", a); return 1; } } while (0)
		if (IS("-dims")) { NEED(); if (!parse_dims(argv[++i], dims)) return 1; }
		else if (IS("-geometry")) { NEED(); geometry = argv[++i]; }
		else if (IS("-iterations")) {
			NEED();
			cd.MaxIterations = (int)strtol(argv[++i], NULL, 0);
			if (cd.MaxIterations <= 0) { printf("ERROR: number of iterations must be > 0.\n"); return 1; }
		}
		else if (IS("-rho-in")) { NEED(); cd.RhoIn = strtod(argv[++i], NULL); }
		else if (IS("-rho-out")) { NEED(); cd.RhoOut = strtod(argv[++i], NULL); }
		else if (IS("-omega")) { NEED(); cd.Omega = strtod(argv[++i], NULL); }
		else if (IS("-x-force")) { NEED(); cd.XForce = strtod(argv[++i], NULL); }
		else if (IS("-verify")) {
			/* preset of src/Main.c:232-265: 16^3 box, periodic in x and y,
			 * omega 1, F = 1e-5, 1000 iterations */
			verify = 1;
			cd.Omega = 1.0; cd.RhoIn = 1.0; cd.RhoOut = 1.0;
			geometry = "box";
			dims[0] = dims[1] = dims[2] = 16;
			cd.XForce = 0.00001;
			cd.MaxIterations = 1000;
			periodic[0] = 1; periodic[1] = 1; periodic[2] = 0;
			printf("#\n# VERIFICATION: verifying flow profile of channel flow.\n#\n");
		}
		else if (IS("-kernel")) { NEED(); kernelName = argv[++i]; }
		else if (IS("-list")) {
			printf("Available kernels to benchmark:\n");
			for (int j = 0; j < g_lbmHipKernelCount; ++j) printf("   %s\n", g_lbmHipKernels[j].Name);
			return 0;
		}
		else if (IS("-t") || IS("-threads")) {
			NEED();
			nThreads = atoi(argv[++i]);
			if (nThreads <= 0) { printf("ERROR: number of threads must be > 0.\n"); return 1; }
		}
		else if (IS("-pin")) { NEED(); ++i; printf("# -pin ignored: the kernels run on the GPU.\n"); }
		else if (IS("-periodic-x")) periodic[0] = 1;
		else if (IS("-periodic-y")) periodic[1] = 1;
		else if (IS("-periodic-z")) periodic[2] = 1;
		else if (IS("-vtk")) {
			/* an optional positive number is the output period (src/Main.c:266-286) */
			cd.VtkOutput = 1;
			if (i + 1 < argc && atoi(argv[i + 1]) > 0) cd.VtkModulus = atoi(argv[++i]);
		}
		else if (IS("-statistics")) {
			NEED();
			cd.StatisticsModulus = atoi(argv[++i]);
			if (cd.StatisticsModulus <= 0) {
				printf("ERROR: the iteration count for -statistics must be > 0.\n");
				return 1;
			}
			statistics = 1;
		}
		else if (IS("-h") || IS("-help")) { usage(); return 1; }
		else if (!strcmp(a, "--")) {
			p.KernelArgs = &argv[i + 1];
			p.nKernelArgs = argc - i - 1;
			break;
		}
		else { printf("ERROR: unknown parameter: %s.\n", a); exit(1); }
#undef IS
#undef NEED
	}

	/* 32-bit PDF indices (src/Main.c:399-406) */
	if ((uint64_t)19 * dims[0] * dims[1] * dims[2] > 0xffffffffull) {
		printf("ERROR: number of PDFs exceed 2^31.\n");
		exit(1);
	}

#ifdef _OPENMP
	extern void omp_set_num_threads(int);
	omp_set_num_threads(nThreads);
#endif

	char devName[256] = "";
	int cus = 0;
	double memClk = 0.0;
	if (LbmHipDeviceInfo(devName, sizeof(devName), &cus, &memClk) == 0)
		printf("# - device:            %s, %d CUs\n", devName, cus);
	printf("# - floating point:    double precision (%zu b)\n", sizeof(PdfT));
	printf("# - iterations:        %d\n", cd.MaxIterations);

	LatticeDesc ld;
	GeoCreateByStr(geometry, dims, periodic, &ld);

	printf("# - geometry:\n");
	printf("#   type:              %s\n", ld.Name);
	printf("#   dimensions:        %d x %d x %d (x, y, z)\n", ld.Dims[0], ld.Dims[1], ld.Dims[2]);
	printf("#   nodes total:       % 10d\n", ld.nObst + ld.nFluid);
	printf("#   nodes fluid:       % 10d (including inlet & outlet)\n", ld.nFluid);
	printf("#   nodes obstacles:   % 10d\n", ld.nObst);
	printf("#   nodes inlet:       % 10d\n", ld.nInlet);
	printf("#   nodes outlet:      % 10d\n", ld.nOutlet);
	printf("#   periodicity:       x: %d y: %d z: %d\n", ld.PeriodicX, ld.PeriodicY, ld.PeriodicZ);
	printf("# - flow:\n");
	printf("#   omega:             %f\n", cd.Omega);
	printf("#   initial density at inlet/outlet:\n");
	printf("#     rho in:          %e\n", cd.RhoIn);
	printf("#     rho out:         %e\n", cd.RhoOut);
	printf("# - host threads:      %d\n", nThreads);

	KernelFunctions * kf = NULL;
	for (int j = 0; j < g_lbmHipKernelCount; ++j)
		if (!strcasecmp(kernelName, g_lbmHipKernels[j].Name)) { kf = &g_lbmHipKernels[j]; break; }
	if (kf == NULL) {
		printf("ERROR: requested kernel \"%s\" not found.\n", kernelName);
		exit(1);
	}
	printf("#\n# - kernel:            %s\n#\n", kf->Name);

	/* -statistics: the kernels write KernelStatistics during the run (the
	 * reference's STATISTICS build) when given the kernel argument */
	char ** kargs = NULL;
	if (statistics) {
		kargs = (char **)malloc(sizeof(char *) * (size_t)(p.nKernelArgs + 1));
		if (!kargs) { printf("ERROR: allocating kernel arguments failed.\n"); return 1; }
		for (int j = 0; j < p.nKernelArgs; ++j) kargs[j] = p.KernelArgs[j];
		kargs[p.nKernelArgs] = (char *)"-statistics";
		p.KernelArgs = kargs;
		p.nKernelArgs += 1;
	}

	KernelData * kd = NULL;
	kf->Init(&ld, &kd, &p);
	if (kd == NULL) { printf("ERROR: kernel initialisation failed.\n"); return 1; }

	if (verify) {
		KernelSetInitialDensity(&ld, kd, &cd);
		KernelSetInitialVelocity(&ld, kd, &cd);
	}

	printf("# starting kernel...\n");
	kd->Kernel(&ld, kd, &cd);

	KernelStatisticsAdv(kd, &ld, &cd, cd.MaxIterations, 1);

	PdfT errorNorm = -1.0;
	KernelVerifiy(&ld, kd, &cd, &errorNorm);

	LbmHipTimings tm;
	memset(&tm, 0, sizeof(tm));
	LbmHipGetTimings(kd, &tm);

	const double duration = kd->Duration;
	const double loopBalance = kd->LoopBalance;
	const double volGiB = loopBalance * ld.nFluid * cd.MaxIterations / 1024. / 1024. / 1024.;

	kf->Deinit(&ld, &kd);
	/* storage tag of the P: line: the reference's PRECISION ("dp"), or the
	 * -prec kernel argument */
	const char * precTag = "dp";
	for (int j = 0; j + 1 < p.nKernelArgs; ++j)
		if (!strcmp(p.KernelArgs[j], "-prec") || !strcmp(p.KernelArgs[j], "--prec")) precTag = p.KernelArgs[j + 1];
	free(kargs);

	const double perf = (double)ld.nFluid * (double)cd.MaxIterations / duration / 1.e6;
	printf("#\n# Evaluation Stats\n");
	printf("#   runtype:     \t%s  \n", "verification");
	printf("#   runtime:           \t%.3f s\n", duration);
	printf("#   iterations:        \t%d  \n", cd.MaxIterations);
	printf("#   fluid cells:       \t%d \n", ld.nFluid);
	printf("# Derived metrics\n");
	printf("#   MEM data vol.:     \t%.2f GByte\n", volGiB);
	printf("#   MEM bandwidth:     \t%.2f GByte/s\n", volGiB / duration);
	printf("#   performance:       \t%.3f MFLUP/s\n", perf);
	if (tm.evenMsAvg > 0.0)   /* kernel argument -launch-events on */
		printf("#   even step:         \t%.4f ms avg  odd step: %.4f ms avg\n", tm.evenMsAvg, tm.oddMsAvg);

	printf("P:   %f MFLUP/s  t: %d  d: %f s  iter: %d  fnodes: %f x1e6  geo: %s  kernel: %s  %s  %s\n",
	       perf, nThreads, duration, cd.MaxIterations, ld.nFluid / 1e6, geometry, kernelName,
	       "VERIFICATION", precTag);

	int rc = 0;
	if (verify) {
		printf("# VERIFICATION: deviation from analytical solution: %e\n", errorNorm);
		if (errorNorm > 0.1) { printf("# VERIFICATION FAILED.\n"); rc = 1; }
		else printf("# VERIFICATION SUCCEEDED.\n");
	}
	GeoFree(&ld);
	return rc;
}
