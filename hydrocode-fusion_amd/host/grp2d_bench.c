/* grp2d_bench.out -- the four-quadrant benchmark window as a plain C program,
 * so that profilers (rocprofv3 --kernel-trace / --pmc) see exactly the launches
 * the benchmark times.
 *
 *   grp2d_bench.out N T_MID STEPS MODE [X_VARIANT Y_VARIANT]
 *
 * Builds the four-quadrant Riemann problem on N x N cells (the states of
 * data_in/two-dim/RP2D_Positive/Config3, split at (0.5, 0.5), gamma 1.4, CFL
 * 0.45, free boundaries), advances it to t ~= T_MID with the reference's solve
 * (whole steps of the frozen tau, no output), then runs STEPS steps with the
 * Newton settings of MODE and prints the device time:
 *   ref      tol = eps, the reference's stopping rule and guess
 *   conv     tol 1e-12, the converge rule, the reference's guess
 *   learned  tol 1e-12, the converge rule, the learned guess (weights from
 *            pstar_mlp.f32, 1313 little-endian floats, next to the library)
 * The window's kernels are the NM_CONV / NM_LEARNED instantiations, whose names
 * differ from those of the approach (NM_REF), so a profile's per-kernel
 * figures for them cover the window alone.  X_VARIANT Y_VARIANT pick the sweep
 * launch forms (hgrp2d_set_sweeps); -1 -1 runs the tuner at T_MID; default 0 0.
 */
#define _POSIX_C_SOURCE 200809L
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "hydrocode_grp.h"

static const double FQ[4][4] = {      /* rho u v p: NE, NW, SW, SE */
    {1.5, 0.0, 0.0, 1.5}, {0.5323, 1.206, 0.0, 0.3}, {0.138, 1.206, 1.206, 0.029}, {0.5323, 0.0, 1.206, 0.3}};

static int load_weights(float *w)
{
    char path[4096];
    const ssize_t len = readlink("/proc/self/exe", path, sizeof path - 32);
    if (len <= 0) return -1;
    path[len] = '\0';
    char *slash = strrchr(path, '/');
    if (!slash) return -1;
    strcpy(slash, "/../lib/pstar_mlp.f32");
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "grp2d_bench: cannot open %s\n", path);
        return -1;
    }
    const size_t got = fread(w, sizeof(float), HGRP_MLP_NP, f);
    fclose(f);
    return got == HGRP_MLP_NP ? 0 : -1;
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s N T_MID STEPS ref|conv|learned [X_VARIANT Y_VARIANT]\n", argv[0]);
        return 2;
    }
    const int N = atoi(argv[1]), steps = atoi(argv[3]);
    const double t_mid = atof(argv[2]);
    const char *mode = argv[4];
    setvbuf(stdout, NULL, _IOLBF, 0);   // progress lines reach a log file as they are printed
    const int xv = argc > 6 ? atoi(argv[5]) : 0, yv = argc > 6 ? atoi(argv[6]) : 0;
    if (N < 2 || N > 65536 || steps < 1 || !(t_mid >= 0.0) ||
        (strcmp(mode, "ref") && strcmp(mode, "conv") && strcmp(mode, "learned"))) {
        fprintf(stderr, "grp2d_bench: bad arguments\n");
        return 2;
    }
    double cfg[HGRP_N_CONF];
    for (int k = 0; k < HGRP_N_CONF; ++k) cfg[k] = INFINITY;
    cfg[0] = 2.0; cfg[1] = 0.3; cfg[4] = 1e-9; cfg[5] = (double)INT_MAX; cfg[6] = 1.4; cfg[7] = 0.45;
    cfg[10] = cfg[11] = 1.0 / N; cfg[13] = cfg[14] = N; cfg[17] = cfg[18] = -4; cfg[33] = 1.0; cfg[41] = 1.9;
    const size_t cells = (size_t)N * N;
    double *q[4];
    for (int f = 0; f < 4; ++f) {
        q[f] = (double *)malloc(cells * sizeof(double));
        if (!q[f]) return 1;
    }
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const int west = (j + 0.5) / N < 0.5, south = (i + 0.5) / N < 0.5;
            const int s = west ? (south ? 2 : 1) : (south ? 3 : 0);
            for (int f = 0; f < 4; ++f) q[f][(size_t)j * N + i] = FQ[s][f];
        }
    hgrp2d *ctx = NULL;
    int rc = hgrp2d_create(&ctx, N, N, cfg);
    if (!rc) rc = hgrp2d_upload(ctx, q[0], q[1], q[2], q[3], NULL);
    double h_s = 0.0;
    if (!rc) rc = hgrp2d_h_over_smax(ctx, &h_s);
    const int k_mid = (int)floor(t_mid / (cfg[7] * h_s));
    double ms = 0.0;
    for (int done = 0; !rc && done < k_mid;) {      // in chunks, with a progress line after each
        const int chunk = k_mid - done < 1000 ? k_mid - done : 1000;
        double part = 0.0;
        rc = hgrp2d_bench(ctx, chunk, 0, &part, NULL, NULL);
        done += chunk;
        ms += part;
        if (!rc) printf("  %d of %d steps to t = %g\n", done, k_mid, t_mid);
    }
    if (rc) {
        fprintf(stderr, "grp2d_bench: setup failed (%d)\n", rc);
        return 1;
    }
    printf("reached t = %.6f after %d steps (reference solve, %.1f ms)\n", k_mid * cfg[7] * h_s, k_mid, ms);
    static float w[HGRP_MLP_NP];
    const int learned = !strcmp(mode, "learned");
    if (learned && load_weights(w)) return 1;
    if (strcmp(mode, "ref"))
        rc = hgrp2d_set_newton(ctx, 1e-12, HGRP_NEWTON_CONVERGE, learned ? w : NULL);
    int tx = xv, ty = yv;
    double tms[2] = {0.0, 0.0};
    if (!rc && (xv < 0 || yv < 0)) rc = hgrp2d_tune(ctx, 4, &tx, &ty, tms);
    else if (!rc) rc = hgrp2d_set_sweeps(ctx, xv, yv);
    if (rc) {
        fprintf(stderr, "grp2d_bench: settings failed (%d)\n", rc);
        return 1;
    }
    printf("mode %s, sweeps X %d (%s) Y %d (%s)\n", mode, tx, hgrp2d_sweep_kernel(ctx, 0), ty,
           hgrp2d_sweep_kernel(ctx, 1));
    rc = hgrp2d_bench(ctx, steps, 0, &ms, NULL, NULL);
    if (rc) {
        fprintf(stderr, "grp2d_bench: window failed (%d)\n", rc);
        return 1;
    }
    printf("%d steps of %dx%d from t = %.6f: %.3f ms, %.4f ms per step, %.4g cell-updates/s\n", steps, N, N,
           k_mid * cfg[7] * h_s, ms, ms / steps, (double)cells * steps / (ms * 1e-3));
    hgrp2d_destroy(ctx);
    for (int f = 0; f < 4; ++f) free(q[f]);
    return 0;
}
