/*
 * hydrocode_2D.c -- the 2D Eulerian program on the GPU library.
 *
 * Same command line, data_in/data_out layout and outputs as the reference's
 * src/hydrocode_2D/hydrocode.c:164-345:
 *
 *     hydrocode_2D.out <in-case> <out-case> <order>[_<scheme>] EUL [n=C ...]
 *
 * run from a directory two levels below data_in/ and data_out/.  config[33]
 * picks the scheme as in the reference: 1 the dimension-split solver, 0 (the
 * default) the unsplit one; both run on the GPU.
 * Exit codes: 1 directory, 2 data, 4 arguments, 5 memory.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrocode_dropin.h"

static double **alloc_rows(int rows, int cols)
{
    double **a = (double **)malloc((size_t)rows * sizeof(double *));
    if (!a)
        return NULL;
    for (int j = 0; j < rows; ++j)
        if (!(a[j] = (double *)malloc((size_t)cols * sizeof(double))))
            return NULL;
    return a;
}

static void free_rows(double **a, int rows)
{
    if (!a)
        return;
    for (int j = 0; j < rows; ++j)
        free(a[j]);
    free(a);
}

int main(int argc, char *argv[])
{
    for (int k = 1; k < N_CONF; k++)
        config[k] = INFINITY;
    char *scheme = NULL;
    arg_preprocess(4, argc, argv, scheme);
    if (argc < 5) {
        printf("NOT appropriate coordinate framework! The framework is missing.\n");
        return 4;
    }
    config[0] = 2.0;

    int N, N_plot;
    double *time_plot;
    struct flu_var FV0 = initialize_2D(argv[1], &N, &N_plot, &time_plot);
    const int n_x = (int)config[13], n_y = (int)config[14];
    const double h_x = config[10], h_y = config[11], gamma = config[6];
    const int order = (int)config[9];
    const int dim_split = config[33] != 0.0;
    int retval = 0;

    struct cell_var_stru *CV = (struct cell_var_stru *)calloc((size_t)N, sizeof(struct cell_var_stru));
    double *cpu_time = (double *)calloc((size_t)N, sizeof(double));
    double **X = alloc_rows(n_x + 1, n_y + 1), **Y = alloc_rows(n_x + 1, n_y + 1);
    if (!CV || !cpu_time || !X || !Y) {
        printf("NOT enough memory!\n");
        return 5;
    }
    for (int k = 0; k < N; ++k) {
        CV[k].RHO = alloc_rows(n_x, n_y); CV[k].U = alloc_rows(n_x, n_y); CV[k].V = alloc_rows(n_x, n_y);
        CV[k].P = alloc_rows(n_x, n_y); CV[k].E = alloc_rows(n_x, n_y);
        if (!CV[k].RHO || !CV[k].U || !CV[k].V || !CV[k].P || !CV[k].E) {
            printf("NOT enough memory! CV[%d]\n", k);
            return 5;
        }
    }
    for (int j = 0; j <= n_x; ++j)
        for (int i = 0; i <= n_y; ++i) {
            X[j][i] = j * h_x;
            Y[j][i] = i * h_y;
        }
    /* file order (line i, column j) -> CV[0].X[j][i]; E as hydrocode.c:245-254 */
    for (int j = 0; j < n_x; ++j)
        for (int i = 0; i < n_y; ++i) {
            const size_t f = (size_t)i * n_x + j;
            CV[0].RHO[j][i] = FV0.RHO[f];
            CV[0].U[j][i] = FV0.U[f];
            CV[0].V[j][i] = FV0.V[f];
            CV[0].P[j][i] = FV0.P[f];
            CV[0].E[j][i] = 0.5 * CV[0].U[j][i] * CV[0].U[j][i] + CV[0].P[j][i] / (gamma - 1.0) / CV[0].RHO[j][i];
            CV[0].E[j][i] += 0.5 * CV[0].V[j][i] * CV[0].V[j][i];
        }

    /* restart at full precision: a FLU_VAR.f64 in the input folder replaces the
     * initial data by its last frame (time starts again at 0) */
    {
        char add_in[FILENAME_MAX + 40], file[FILENAME_MAX + 60];
        example_io(argv[1], add_in, 1);
        snprintf(file, sizeof file, "%sFLU_VAR.f64", add_in);
        int rx, ry, rn;
        double *rtp = NULL, *rfr = NULL;
        const int rr = file_2D_read_f64(file, &rx, &ry, &rn, &rtp, &rfr);
        if (rr == 0 && (rx != n_x || ry != n_y)) {
            printf("FLU_VAR.f64 is %d x %d, the case is %d x %d: not used.\n", rx, ry, n_x, n_y);
        } else if (rr == 0) {
            const size_t cells = (size_t)n_x * n_y;
            const double *last = rfr + (size_t)(rn - 1) * 5 * cells;
            double **dst[5] = {CV[0].RHO, CV[0].U, CV[0].V, CV[0].P, CV[0].E};
            for (int q = 0; q < 5; ++q)
                for (int j = 0; j < n_x; ++j)
                    memcpy(dst[q][j], last + q * cells + (size_t)j * n_y, (size_t)n_y * sizeof(double));
            printf("Restart from %s (frame time %.10g).\n", file, rtp[rn - 1]);
        }
        free(rtp);
        free(rfr);
    }

    if (strcmp(argv[4], "EUL") != 0) {
        printf("NOT appropriate coordinate framework! The framework is %s.\n", argv[4]);
        retval = 4;
        goto cleanup;
    }
    config[8] = 0.0;
    if (order == 1)
        config[41] = 0.0;          /* first order: alpha = 0, still the GRP solver */
    if (order != 1 && order != 2) {
        printf("NOT appropriate order of the scheme! The order is %d.\n", order);
        retval = 4;
        goto cleanup;
    }
    if (dim_split)       /* hydrocode.c:264-267 */
        GRP_solver_2D_split_EUL_source(n_x, n_y, CV, X, Y, cpu_time, argv[2], N, &N_plot, time_plot);
    else
        GRP_solver_2D_EUL_source(n_x, n_y, CV, X, Y, cpu_time, argv[2], N, &N_plot, time_plot);

#ifndef NODATPLOT
    file_2D_write(n_x, n_y, N_plot, CV, X, Y, cpu_time, argv[2], time_plot);
#endif
    {
        char add_out[FILENAME_MAX + 40], file[FILENAME_MAX + 60];
        example_io(argv[2], add_out, 0);
        snprintf(file, sizeof file, "%sFLU_VAR.f64", add_out);
        if (file_2D_write_f64(file, n_x, n_y, N_plot, CV, time_plot))
            retval = 1;
    }
#ifndef NOTECPLOT
    file_2D_write_POINT_TEC(n_x, n_y, 1, CV + N_plot - 1, X, Y, cpu_time, argv[2], time_plot + N_plot - 1);
#endif

cleanup:
    free(FV0.RHO); free(FV0.U); free(FV0.V); free(FV0.P);
    for (int k = 0; k < N; ++k) {
        free_rows(CV[k].RHO, n_x); free_rows(CV[k].U, n_x); free_rows(CV[k].V, n_x);
        free_rows(CV[k].P, n_x); free_rows(CV[k].E, n_x);
    }
    free(CV);
    free_rows(X, n_x + 1);
    free_rows(Y, n_x + 1);
    free(cpu_time);
    free(time_plot);
    return retval;
}
