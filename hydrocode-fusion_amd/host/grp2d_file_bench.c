/* grp2d_file_bench.out -- a 2D workload read from a file, run as a plain C
 * program so that profilers (rocprofv3 --kernel-trace / --pmc) see exactly
 * the launches of the benchmark's unsplit and two-fluid legs.
 *
 *   grp2d_file_bench.out FILE STEPS ref|conv|learned [X_VARIANT Y_VARIANT | g2d]
 *
 * FILE, little-endian (written by tests/cases.py: write_grid_bin):
 *   int32 m, n; float64 config[400] (the reference's config[]);
 *   float64 rho[m*n], u[m*n], v[m*n], p[m*n] (cell (j, i) at j*n + i);
 *   int32 two_fluid; if 1: float64 gamma_b, float64 frac_a[m*n].
 * Uploads the data (and the volume fraction), sets the Newton mode (ref: the
 * reference's solve; conv: tol 1e-12 and the converge rule; learned: the same
 * from the learned guess, weights from pstar_mlp.f32 next to the library) and
 * runs STEPS steps from t = 0 of the scheme config[33] selects (split: the
 * X(tau/2) Y(tau) X(tau/2) benchmark form, launch variants as given, default
 * 0 0; unsplit: one step launch per tau, with GRP_2D_flux's alternative
 * interface solver G2D when the last argument is "g2d"), then prints the device
 * time.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "hydrocode_grp.h"

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
        fprintf(stderr, "grp2d_file_bench: cannot open %s\n", path);
        return -1;
    }
    const size_t got = fread(w, sizeof(float), HGRP_MLP_NP, f);
    fclose(f);
    return got == HGRP_MLP_NP ? 0 : -1;
}

static int read_all(FILE *f, void *p, size_t size, size_t n) { return fread(p, size, n, f) == n ? 0 : -1; }

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s FILE STEPS ref|conv|learned [X_VARIANT Y_VARIANT | g2d]\n", argv[0]);
        return 2;
    }
    setvbuf(stdout, NULL, _IOLBF, 0);
    const int steps = atoi(argv[2]);
    const char *mode = argv[3];
    const int xv = argc > 5 ? atoi(argv[4]) : 0, yv = argc > 5 ? atoi(argv[5]) : 0;
    const int g2d = argc == 5 && !strcmp(argv[4], "g2d");
    if (steps < 1 || (strcmp(mode, "ref") && strcmp(mode, "conv") && strcmp(mode, "learned"))) {
        fprintf(stderr, "grp2d_file_bench: bad arguments\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "grp2d_file_bench: cannot open %s\n", argv[1]);
        return 2;
    }
    int mn[2];
    static double cfg[HGRP_N_CONF];
    if (read_all(f, mn, 4, 2) || mn[0] < 1 || mn[1] < 1 || mn[0] > 65536 || mn[1] > 65536 ||
        read_all(f, cfg, 8, HGRP_N_CONF)) {
        fprintf(stderr, "grp2d_file_bench: bad header\n");
        return 2;
    }
    const int m = mn[0], n = mn[1];
    const size_t cells = (size_t)m * n;
    double *q[5] = {NULL, NULL, NULL, NULL, NULL};
    for (int k = 0; k < 4; ++k) {
        q[k] = (double *)malloc(cells * sizeof(double));
        if (!q[k] || read_all(f, q[k], 8, cells)) {
            fprintf(stderr, "grp2d_file_bench: short file\n");
            return 2;
        }
    }
    int two = 0;
    double gamma_b = 0.0;
    if (read_all(f, &two, 4, 1)) two = 0;
    if (two) {
        q[4] = (double *)malloc(cells * sizeof(double));
        if (!q[4] || read_all(f, &gamma_b, 8, 1) || read_all(f, q[4], 8, cells)) {
            fprintf(stderr, "grp2d_file_bench: short file (volume fraction)\n");
            return 2;
        }
    }
    fclose(f);
    const int unsplit = cfg[33] == 0.0;
    hgrp2d *ctx = NULL;
    int rc = hgrp2d_create(&ctx, m, n, cfg);
    if (!rc) rc = hgrp2d_upload(ctx, q[0], q[1], q[2], q[3], NULL);
    if (!rc && two) rc = hgrp2d_set_fluids(ctx, gamma_b, q[4]);
    static float w[HGRP_MLP_NP];
    const int learned = !strcmp(mode, "learned");
    if (!rc && learned && load_weights(w)) rc = -1;
    if (!rc && strcmp(mode, "ref")) rc = hgrp2d_set_newton(ctx, 1e-12, HGRP_NEWTON_CONVERGE, learned ? w : NULL);
    if (!rc && !unsplit && !two) rc = hgrp2d_set_sweeps(ctx, xv, yv);
    if (!rc && g2d) rc = hgrp2d_set_interface_solver(ctx, HGRP_SOLVER_G2D);
    if (rc) {
        fprintf(stderr, "grp2d_file_bench: setup failed (%d)\n", rc);
        return 1;
    }
    printf("%d x %d, %s, mode %s, kernels %s / %s\n", m, n, unsplit ? "unsplit" : two ? "split, two fluids" : "split",
           mode, hgrp2d_sweep_kernel(ctx, unsplit ? 2 : 0), unsplit ? "-" : hgrp2d_sweep_kernel(ctx, 1));
    double ms = 0.0;
    rc = unsplit ? hgrp2d_bench_unsplit(ctx, steps, 0, &ms, NULL) : hgrp2d_bench(ctx, steps, 0, &ms, NULL, NULL);
    if (rc) {
        fprintf(stderr, "grp2d_file_bench: run failed (%d)\n", rc);
        return 1;
    }
    printf("%d steps from t = 0: %.3f ms, %.4f ms per step, %.4g cell-updates/s\n", steps, ms, ms / steps,
           (double)cells * steps / (ms * 1e-3));
    hgrp2d_destroy(ctx);
    for (int k = 0; k < 5; ++k) free(q[k]);
    return 0;
}
