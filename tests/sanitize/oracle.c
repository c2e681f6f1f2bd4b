/* The oracle's list build, list and aa runs and their shifted-storage
 * variants on five small geometries; built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_sanitizers.py. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lbm_oracle.h"

int main(void)
{
	const int types[] = {0, 1, 2, 3, 4};   /* box channel pipe blocks fluid */
	for (int ti = 0; ti < 5; ++ti) {
		int dims[3] = {18, 14, 12}, per[3] = {ti == 0, ti == 0, 0}, pf[3];
		const size_t N = (size_t)dims[0] * dims[1] * dims[2];
		int * mask = calloc(N, sizeof(int));
		orc_geometry(types[ti], 3, dims, per, mask, pf);
		uint32_t * grid = calloc(N, 4), * coords = calloc(3 * N, 4);
		int blk[3] = {ti == 3 ? 4 : 0, ti == 3 ? 4 : 0, ti == 3 ? 4 : 0};
		const int nF = orc_list_number(mask, dims, blk, grid, coords);
		const int nC = orc_pad_cells(nF, 8, NULL);
		uint32_t * adj = calloc((size_t)nF * 18 + 1, 4);
		orc_list_adjacency(mask, dims, pf, grid, coords, nF, nC, adj);
		const int runs = orc_list_runs(adj, nF, 3, NULL);
		double * pdf = calloc((size_t)19 * nC, sizeof(double));
		orc_list_t L = {{dims[0], dims[1], dims[2]}, nF, nC, mask, grid, adj, pdf, -1};
		double w[19];
		for (int d = 0; d < 19; ++d) w[d] = ORC_W[d];
		for (int z = 0; z < dims[2]; ++z) for (int y = 0; y < dims[1]; ++y) for (int x = 0; x < dims[0]; ++x)
			if (mask[((size_t)z * dims[1] + y) * dims[0] + x]) orc_list_set_node(&L, x, y, z, w);
		orc_list_run(&L, 20, 1.2, 3.0 / 16, 1e-5, 1, 3, 4);
		orc_list_run(&L, 20, 1.2, 3.0 / 16, 1e-5, 1, 1, 1);
		double * dump = calloc((size_t)nF * 19 + 1, sizeof(double));
		orc_list_dump(&L, dump);
		float * st = calloc((size_t)19 * nC, sizeof(float));
		for (int prec = ORC_PREC_FP32S; prec <= ORC_PREC_FP16S; ++prec) {
			orc_shifted_from_nodes(&L, st, prec, 12, dump);
			orc_list_run_shifted(&L, st, prec, 12, 20, 1.2, 3.0 / 16, 1e-5, 1);
			orc_shifted_to_nodes(&L, st, prec, 12, dump);
		}
		orc_aa_t A;
		orc_aa_init(&A, mask, dims, pf);
		for (int z = 0; z < dims[2]; ++z) for (int y = 0; y < dims[1]; ++y) for (int x = 0; x < dims[0]; ++x)
			if (mask[((size_t)z * dims[1] + y) * dims[0] + x]) orc_aa_set_node(&A, x, y, z, w);
		orc_aa_run(&A, 20, 1.2, 3.0 / 16, 1e-5, 1);
		orc_aa_dump(&A, dump);
		const size_t nAa = (size_t)19 * (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2);
		float * sa = calloc(nAa, sizeof(float));
		for (int prec = ORC_PREC_FP32S; prec <= ORC_PREC_FP16S; ++prec) {
			orc_aa_shifted_from_nodes(&A, sa, prec, 12, dump);
			if (orc_aa_run_shifted(&A, sa, prec, 12, 20, 1.2, 3.0 / 16, 1e-5, 1)) return 2;
			orc_aa_shifted_to_nodes(&A, sa, prec, 12, dump);
		}
		printf("type %d: nFluid %d nCells %d runs %d bb %d dump[0] %f\n", ti, nF, nC, runs, A.nBB, nF ? dump[0] : 0.0);
		orc_aa_free(&A);
		free(sa); free(st); free(dump); free(pdf); free(adj); free(grid); free(coords); free(mask);
	}
	return 0;
}
