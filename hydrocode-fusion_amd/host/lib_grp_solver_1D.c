/*
 * lib_grp_solver_1D.c -- GRP_solver_EUL_source (the 1D Eulerian GRP scheme)
 * on the GPU.
 *
 * Drop-in for src/finite_volume/grp_solver_EUL_source.c:27-334: same
 * signature, same use of config[] (reads 1, 4, 5, 6, 7, 10, 16, 17, 41;
 * writes config[5] = k on exit), same frames CV.X[nt], N_plot and time_plot
 * results.  The time loop is one instance of the batched 1D engine
 * (hgrp1d_*, include/hydrocode_grp.h), which stops at the start of the step
 * where the time first reaches the next plot time; the state is then copied
 * into the current frame and the loop continues in the next one, as the
 * reference's frame copy (:116-125) does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hydrocode_dropin.h"
#include "hydrocode_grp.h"

static const char *bc_name_1d(int bc)
{
    switch (bc) {
    case -1: return "Initial";
    case -2: return "Reflective";
    case -4: return "Free";
    case -7: return "Periodic";
    case -24: return "Reflective + Free";
    default: return NULL;
    }
}

static int download_frame(hgrp1d *ctx, const struct cell_var_stru *CV, int nt)
{
    return hgrp1d_download(ctx, CV->RHO[nt], CV->U[nt], CV->P[nt], CV->E[nt]);
}

void GRP_solver_EUL_source(const int m, struct cell_var_stru CV, double *cpu_time, int *N_plot, double time_plot[])
{
    const double t_all = config[1];
    const double h = config[10];
    hgrp1d *ctx = NULL;
    int k = 0, nt = 0, status = 0;
    double time_c = 0.0;
    const clock_t tic = clock();

    const char *name = bc_name_1d((int)config[17]);
    if (name && (int)config[17] == -1)
        printf("%s boudary conditions in x direction at time %g .\n", name, 0.0);
    else if (name)
        printf("%s boudary conditions in x direction.\n", name);

    int rc = hgrp1d_create(&ctx, 1, &m, &h, config);
    if (rc) {
        printf("GPU solver set-up failed (%d)\n", rc);
        goto done;
    }
    rc = hgrp1d_upload(ctx, CV.RHO[0], CV.U[0], CV.P[0], CV.E[0]);
    if (rc) {
        printf("GPU upload failed (%d)\n", rc);
        goto done;
    }
    /* the test at the start of step 1 (time 0): frame 0 keeps the initial data */
    if (0.0 >= time_plot[0] && nt < *N_plot - 1) {
        memcpy(CV.RHO[1], CV.RHO[0], (size_t)m * sizeof(double));
        memcpy(CV.U[1], CV.U[0], (size_t)m * sizeof(double));
        memcpy(CV.P[1], CV.P[0], (size_t)m * sizeof(double));
        memcpy(CV.E[1], CV.E[0], (size_t)m * sizeof(double));
        nt = 1;
    }
    for (;;) {
        const double t_pause = nt < *N_plot - 1 ? time_plot[nt] : INFINITY;
        int paused = 0;
        rc = hgrp1d_run_until(ctx, 1, &t_pause, &paused);
        if (rc) {
            printf("GPU time loop failed (%d)\n", rc);
            goto done;
        }
        if (!paused)
            break;
        /* the frame copy: frame nt keeps this state, the loop goes on in frame nt + 1 */
        if (download_frame(ctx, &CV, nt)) {
            printf("GPU download failed\n");
            goto done;
        }
        if (cpu_time)
            cpu_time[nt] = ((double)clock() - (double)tic) / (double)CLOCKS_PER_SEC;
        ++nt;
    }
    rc = hgrp1d_results(ctx, &k, &status, &time_c);
    if (rc) {
        printf("GPU results failed (%d)\n", rc);
        k = 0;
        time_c = 0.0;
        goto done;
    }
    if (download_frame(ctx, &CV, nt))
        printf("GPU download failed\n");
    {
        const double sec = ((double)clock() - (double)tic) / (double)CLOCKS_PER_SEC;
        if (cpu_time)
            cpu_time[nt] = sec;
        if (status != 1) {
            printf("\nTime is up at time step %d.\n", k);
            printf("The cost of CPU time for 1D-GRP Eulerian scheme for this problem is %g seconds.\n", sec);
        }
    }

done:
    /* the reference's return_NULL epilogue (:286-294) */
    config[5] = (double)k;
    *N_plot = nt + 1;
    if (isfinite(time_c))
        time_plot[nt] = time_c;
    else if (isfinite(t_all))
        time_plot[nt] = t_all;
    else if (isfinite(config[16]))
        time_plot[nt] = k * config[16];
    if (ctx)
        hgrp1d_destroy(ctx);
}
