/* grp1d_bench.out -- the batched 1D ensemble run as a plain C program, so that
 * profilers (rocprofv3 --kernel-trace / --pmc) see the launches the
 * benchmark times without a Python interpreter in the process.
 *
 *   grp1d_bench.out FILE [GRAPH [BS_IFACE BS_UPDATE]]
 *
 * FILE holds the ensemble, little-endian: int32 n_inst, int32 cells[n_inst],
 * then float64 rho[], u[], p[] over all cells (instances concatenated; cell
 * size h = 1 / cells of the instance; domain [0, 1]).  Gamma 1.4, CFL 0.45,
 * free boundaries, t_end 0.15, as the benchmark's ensemble leg.  GRAPH 1
 * (default) runs captured graphs of 32 steps, 0 plain launches.  Block sizes:
 * BS_IFACE BS_UPDATE, or the autotuner (hgrp1d_tune, 32 probe steps) when
 * they are not given.  Prints the device time and the cell-updates per second.
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hydrocode_grp.h"

static int read_all(FILE *f, void *p, size_t size, size_t n) { return fread(p, size, n, f) == n ? 0 : -1; }

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s FILE [GRAPH [BS_IFACE BS_UPDATE]]\n", argv[0]);
        return 2;
    }
    const int graph = argc > 2 ? atoi(argv[2]) : 1;
    setvbuf(stdout, NULL, _IOLBF, 0);
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "grp1d_bench: cannot open %s\n", argv[1]);
        return 2;
    }
    int n_inst = 0;
    if (read_all(f, &n_inst, 4, 1) || n_inst < 1 || n_inst > (1 << 20)) {
        fprintf(stderr, "grp1d_bench: bad header\n");
        return 2;
    }
    int *cells = (int *)malloc((size_t)n_inst * sizeof(int));
    double *h = (double *)malloc((size_t)n_inst * sizeof(double));
    if (!cells || !h || read_all(f, cells, 4, (size_t)n_inst)) return 2;
    size_t total = 0;
    for (int i = 0; i < n_inst; ++i) {
        if (cells[i] < 1) return 2;
        total += (size_t)cells[i];
        h[i] = 1.0 / cells[i];
    }
    double *q[3];
    for (int k = 0; k < 3; ++k) {
        q[k] = (double *)malloc(total * sizeof(double));
        if (!q[k] || read_all(f, q[k], 8, total)) {
            fprintf(stderr, "grp1d_bench: short file\n");
            return 2;
        }
    }
    fclose(f);
    double cfg[HGRP_N_CONF];
    for (int k = 0; k < HGRP_N_CONF; ++k) cfg[k] = INFINITY;
    cfg[0] = 1.0; cfg[1] = 0.15; cfg[4] = 1e-9; cfg[5] = (double)INT_MAX; cfg[6] = 1.4; cfg[7] = 0.45;
    cfg[17] = -4; cfg[41] = 1.9;
    hgrp1d *ctx = NULL;
    int rc = hgrp1d_create(&ctx, n_inst, cells, h, cfg);
    if (!rc) rc = hgrp1d_upload(ctx, q[0], q[1], q[2], NULL);
    int bi = argc > 4 ? atoi(argv[3]) : 0, bu = argc > 4 ? atoi(argv[4]) : 0;
    if (!rc && argc <= 4) rc = hgrp1d_tune(ctx, 32, &bi, &bu);
    if (!rc) printf("block sizes %d/%d; running\n", bi, bu);
    if (!rc && argc > 4) rc = hgrp1d_set_blocks(ctx, bi, bu);
    double ms = 0.0;
    long long cu = 0;
    if (!rc) rc = hgrp1d_run(ctx, graph, &ms, &cu);
    if (rc) {
        fprintf(stderr, "grp1d_bench: failed (%d)\n", rc);
        return 1;
    }
    printf("%d instances, %zu cells, block sizes %d/%d, %s: %.3f ms, %lld cell-updates, %.4g cell-updates/s\n", n_inst,
           total, bi, bu, graph ? "graphs" : "plain launches", ms, cu, (double)cu / (ms * 1e-3));
    hgrp1d_destroy(ctx);
    return 0;
}
