/*
 * lib_grp_solver_2D.c -- GRP_solver_2D_split_EUL_source and
 * GRP_solver_2D_EUL_source on the GPU.
 *
 * Drop-ins for src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379 and
 * src/finite_volume/grp_solver_2D_EUL_source.c:75-330: same signatures, same
 * use of config[] (reads 1, 4, 5, 6, 7, 10, 11, 16, 17, 18, 41; writes
 * config[5] = k on exit), same CV frames, N_plot and time_plot results and the
 * same Tecplot frames at plot times.  The time loop runs on one GPU through
 * the C ABI of hydrocode_grp.h; this file only moves the caller's row pointers
 * in and out.  The scratch members of CV[0] (slopes, fluxes, interface values)
 * are never allocated and stay NULL, as the reference leaves them on return.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrocode_dropin.h"
#include "hydrocode_grp.h"

struct frame_ctx {
    int m, n;
    struct cell_var_stru *CV;
    double **X, **Y;
    double *cpu_time;
    const char *problem;
    double *time_plot;
    hgrp2d *ctx;
    double *buf;            /* 5 * m * n contiguous doubles */
};

static void pack(const struct frame_ctx *f, const struct cell_var_stru *cv)
{
    const size_t mn = (size_t)f->m * f->n;
    double **src[5] = {cv->RHO, cv->U, cv->V, cv->P, cv->E};
    for (int q = 0; q < 5; ++q)
        for (int j = 0; j < f->m; ++j)
            memcpy(f->buf + q * mn + (size_t)j * f->n, src[q][j], (size_t)f->n * sizeof(double));
}

static int unpack_device_state(const struct frame_ctx *f, struct cell_var_stru *cv)
{
    const size_t mn = (size_t)f->m * f->n;
    int rc = hgrp2d_download(f->ctx, f->buf, f->buf + mn, f->buf + 2 * mn, f->buf + 3 * mn, f->buf + 4 * mn);
    if (rc)
        return rc;
    double **dst[5] = {cv->RHO, cv->U, cv->V, cv->P, cv->E};
    for (int q = 0; q < 5; ++q)
        for (int j = 0; j < f->m; ++j)
            memcpy(dst[q][j], f->buf + q * mn + (size_t)j * f->n, (size_t)f->n * sizeof(double));
    return 0;
}

/* the reference writes the Tecplot frame of CV[nt] at time_plot[nt_plot] (:157-162) */
static int on_frame(void *user, int nt, int nt_plot, double time_c)
{
    (void)time_c;
    struct frame_ctx *f = (struct frame_ctx *)user;
    if (unpack_device_state(f, f->CV + nt))
        return 1;
#ifndef NOTECPLOT
    file_2D_write_POINT_TEC(f->m, f->n, 1, f->CV + nt, f->X, f->Y, f->cpu_time, f->problem, f->time_plot + nt_plot);
#endif
    return 0;
}

static const char *bc_name(int bc)
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

/* split = 1: the dimension-split scheme, 0: the unsplit one (config[33] of the context) */
static void run_2d(const int m, const int n, struct cell_var_stru *CV, double **X, double **Y, double *cpu_time,
                   const char *problem, int N_T, int *N_plot, double time_plot[], int split)
{
    char dev[256];
    hgrp_device_name(dev, (int)sizeof dev);
    printf("@@ GPU device: %s\n", dev[0] ? dev : "(none)");
    for (int d = 0; d < 2; ++d) {
        const char *name = bc_name((int)config[17 + d]);
        if (name)
            printf("%s boudary conditions in %c direction.\n", name, d ? 'y' : 'x');
    }

    struct frame_ctx f = {m, n, CV, X, Y, cpu_time, problem, time_plot, NULL, NULL};
    hgrp2d_result res = {0, 0, 1, 0, 0.0, config[16], 0.0};
    int rc = 0;
    f.buf = (double *)malloc((size_t)5 * m * n * sizeof(double));
    if (f.buf == NULL) {
        printf("NOT enough memory! host staging buffer\n");
        goto done;
    }
    {
        double cfg[HGRP_N_CONF];
        memcpy(cfg, config, sizeof cfg);
        cfg[33] = split ? 1.0 : 0.0;
        rc = hgrp2d_create(&f.ctx, m, n, cfg);
    }
    if (rc) {
        printf("GPU solver set-up failed (%d)\n", rc);
        goto done;
    }
    pack(&f, CV);
    const size_t mn = (size_t)m * n;
    rc = hgrp2d_upload(f.ctx, f.buf, f.buf + mn, f.buf + 2 * mn, f.buf + 3 * mn, f.buf + 4 * mn);
    if (rc) {
        printf("GPU upload failed (%d)\n", rc);
        goto done;
    }
    rc = hgrp2d_solve(f.ctx, time_plot, *N_plot, N_T, on_frame, &f, &res);
    if (rc) {
        printf("GPU time loop failed (%d)\n", rc);
        res.k = 0; res.nt = 0; res.time_c = 0.0;
        goto done;
    }
    if (unpack_device_state(&f, CV + res.nt))
        printf("GPU download failed\n");
    if (cpu_time)
        cpu_time[res.nt] = res.seconds;

done:
    /* the reference's return_NULL epilogue (:333-341) */
    config[5] = (double)res.k;
    *N_plot = res.nt + 1;
    if (isfinite(res.time_c))
        time_plot[res.nt] = res.time_c;
    else if (isfinite(config[1]))
        time_plot[res.nt] = config[1];
    else if (isfinite(res.tau))
        time_plot[res.nt] = res.k * res.tau;
    if (f.ctx)
        hgrp2d_destroy(f.ctx);
    free(f.buf);
}

void GRP_solver_2D_split_EUL_source(const int m, const int n, struct cell_var_stru *CV, double **X, double **Y,
                                    double *cpu_time, const char *problem, int N_T, int *N_plot,
                                    double time_plot[])
{
    run_2d(m, n, CV, X, Y, cpu_time, problem, N_T, N_plot, time_plot, 1);
}

void GRP_solver_2D_EUL_source(const int m, const int n, struct cell_var_stru *CV, double **X, double **Y,
                              double *cpu_time, const char *problem, int N_T, int *N_plot, double time_plot[])
{
    run_2d(m, n, CV, X, Y, cpu_time, problem, N_T, N_plot, time_plot, 0);
}
