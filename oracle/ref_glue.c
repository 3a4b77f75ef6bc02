/*
 * ref_glue.c -- caller-side glue for the reference build in oracle/_ref.
 * TEST INFRASTRUCTURE ONLY: linked with the reference's own sources (compiled
 * from /root/reference/src by oracle/Makefile) to produce golden vectors and to
 * cross-check the C restatement.  It plays the part of the reference's main
 * programs: it owns the global config[] (as src/hydrocode_2D/hydrocode.c:123
 * does) and hands flat arrays to the reference's entry points.
 *
 * Interface-side vectors use 15 doubles in the order
 *   gamma, lambda_u, lambda_v, rho, u, v, p, d_rho, d_u, d_v, d_p, t_rho, t_u, t_v, t_p
 * (the same order as orc_ifv in grp_oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>

#include "var_struc.h"
#include "riemann_solver.h"
/* The reference's Toro solver file, compiled in this unit so that its static
 * GuessP can be called (it is no part of the 2D path otherwise). */
#include "riemann_solver_starPU.c"
#include "flux_calc.h"
#include "inter_process.h"
#include "finite_volume.h"

double config[N_CONF];

void ref_config_reset(void)
{
    for (int k = 1; k < N_CONF; ++k) config[k] = INFINITY;
    config[0] = 2.0;
}
void ref_config_set(int idx, double val) { config[idx] = val; }
double ref_config_get(int idx) { return config[idx]; }

static void to_ifv(struct i_f_var *f, const double *s)
{
    memset(f, 0, sizeof *f);
    f->gamma = s[0]; f->lambda_u = s[1]; f->lambda_v = s[2];
    f->RHO = s[3]; f->U = s[4]; f->V = s[5]; f->P = s[6];
    f->d_rho = s[7]; f->d_u = s[8]; f->d_v = s[9]; f->d_p = s[10];
    f->t_rho = s[11]; f->t_u = s[12]; f->t_v = s[13]; f->t_p = s[14];
}

/* in: gammaL gammaR u_L u_R p_L p_R c_L c_R eps tol ; N
 * out: U_star P_star CRW0 CRW1 gap */
void ref_riemann_exact(const double *in, int N, double *out)
{
    _Bool CRW[2] = {false, false};
    double us = 0, ps = 0;
    double gap = Riemann_solver_exact(&us, &ps, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7],
                                      CRW, in[8], in[9], N);
    out[0] = us; out[1] = ps; out[2] = CRW[0]; out[3] = CRW[1]; out[4] = gap;
}

/* in: gamma u_L u_R p_L p_R c_L c_R eps tol ; N ; out as above */
void ref_riemann_exact_ben(const double *in, int N, double *out)
{
    _Bool CRW[2] = {false, false};
    double us = 0, ps = 0;
    double gap = Riemann_solver_exact_Ben(&us, &ps, in[0], in[1], in[2], in[3], in[4], in[5], in[6],
                                          CRW, in[7], in[8], N);
    out[0] = us; out[1] = ps; out[2] = CRW[0]; out[3] = CRW[1]; out[4] = gap;
}

/* out: wave_speed[2], D[0..3], U[0..3], U_star[6]  (16 doubles) */
void ref_q1d(const double *Ls, const double *Rs, double eps, double atc, double *out)
{
    struct i_f_var L, R;
    to_ifv(&L, Ls); to_ifv(&R, Rs);
    double ws[2], D[6], U[6], Us[6];
    linear_GRP_solver_Edir_Q1D(ws, D, U, Us, &L, &R, eps, atc);
    out[0] = ws[0]; out[1] = ws[1];
    for (int q = 0; q < 4; ++q) { out[2 + q] = D[q]; out[6 + q] = U[q]; }
    for (int q = 0; q < 6; ++q) out[10 + q] = Us[q];
}

/* linear_GRP_solver_Edir_G2D (single-fluid build: z_a and phi are the
 * solver's constants).  out[20]: wave_speed[2], D[6], U[6], U_star[6]. */
void ref_g2d(const double *Ls, const double *Rs, double eps, double atc, double *out)
{
    struct i_f_var L, R;
    to_ifv(&L, Ls); to_ifv(&R, Rs);
    linear_GRP_solver_Edir_G2D(out, out + 2, out + 8, out + 14, &L, &R, eps, atc);
}

/* GRP_2D_flux; config[4] (eps) must be set.  out: F_rho F_u F_v F_e RHO_int U_int V_int P_int */
int ref_grp_2d_flux(const double *Ls, const double *Rs, double n_x, double n_y, double tau, double *out)
{
    struct i_f_var L, R;
    to_ifv(&L, Ls); to_ifv(&R, Rs);
    L.n_x = n_x; L.n_y = n_y; R.n_x = n_x; R.n_y = n_y;
    int rc = GRP_2D_flux(&L, &R, tau);
    out[0] = L.F_rho; out[1] = L.F_u; out[2] = L.F_v; out[3] = L.F_e;
    out[4] = L.RHO_int; out[5] = L.U_int; out[6] = L.V_int; out[7] = L.P_int;
    return rc;
}

/* The members of struct i_f_var that a single-fluid build has, in their
 * declaration order (src/include/var_struc.h:116-131): 34 doubles. */
#define IFV_MEMBERS(X) X(n_x) X(n_y) X(delta_x) X(delta_y) X(length) X(lambda_u) X(lambda_v) X(gamma) \
    X(RHO) X(P) X(U) X(V) X(RHO_int) X(P_int) X(U_int) X(V_int) X(F_rho) X(F_e) X(F_u) X(F_v)          \
    X(U_rho) X(U_e) X(U_u) X(U_v) X(d_rho) X(d_e) X(d_u) X(d_v) X(d_p) X(t_rho) X(t_e) X(t_u) X(t_v) X(t_p)

static void ifv_from_members(struct i_f_var *f, const double *a)
{
    int k = 0;
    memset(f, 0, sizeof *f);
#define SET(name) f->name = a[k++];
    IFV_MEMBERS(SET)
#undef SET
}

static void ifv_to_members(double *a, const struct i_f_var *f)
{
    int k = 0;
#define GET(name) a[k++] = f->name;
    IFV_MEMBERS(GET)
#undef GET
}

/* GRP_2D_flux on whole structs (34 members each, in and out): its side effects
 * on both sides as well as its results.  config[4] (eps) must be set. */
int ref_grp_2d_flux_struct(const double *Lm, const double *Rm, double tau, double *Lout, double *Rout)
{
    struct i_f_var L, R;
    ifv_from_members(&L, Lm);
    ifv_from_members(&R, Rm);
    int rc = GRP_2D_flux(&L, &R, tau);
    ifv_to_members(Lout, &L);
    ifv_to_members(Rout, &R);
    return rc;
}

/* GuessP of riemann_solver_starPU.c:20-58.  in: DL DR UL UR PL PR CL CR GammaL GammaR eps */
double ref_guessp(const double *in)
{
    return GuessP(in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8], in[9], in[10]);
}

/* out: D[3], U[3] */
void ref_grp_edir(const double *Ls, const double *Rs, double eps, double atc, double *out)
{
    struct i_f_var L, R;
    to_ifv(&L, Ls); to_ifv(&R, Rs);
    double D[3], U[3];
    linear_GRP_solver_Edir(D, U, &L, &R, eps, atc);
    for (int q = 0; q < 3; ++q) { out[q] = D[q]; out[3 + q] = U[q]; }
}

/* Runs one of the 2D solvers.  fields: 5*N_T blocks of m*n doubles, order
 * rho u v p e, frame-major ([f][k][j*n+i]); frame 0 = initial data.  config[]
 * must be prepared by the caller (ref_config_*). */
static void ref_2d(int m, int n, double *fields, int N_T, int *N_plot, double *time_plot, int split)
{
    struct cell_var_stru *CV = calloc((size_t)N_T, sizeof *CV);
    double **rows = malloc((size_t)5 * N_T * m * sizeof(double *));
    const size_t mn = (size_t)m * n;
    for (int f = 0; f < 5; ++f)
        for (int k = 0; k < N_T; ++k) {
            double **r = rows + ((size_t)f * N_T + k) * m;
            for (int j = 0; j < m; ++j) r[j] = fields + ((size_t)f * N_T + k) * mn + (size_t)j * n;
            switch (f) {
            case 0: CV[k].RHO = r; break;
            case 1: CV[k].U = r; break;
            case 2: CV[k].V = r; break;
            case 3: CV[k].P = r; break;
            default: CV[k].E = r; break;
            }
        }
    double **X = malloc((size_t)(m + 1) * sizeof(double *)), **Y = malloc((size_t)(m + 1) * sizeof(double *));
    for (int j = 0; j <= m; ++j) {
        X[j] = malloc((size_t)(n + 1) * sizeof(double));
        Y[j] = malloc((size_t)(n + 1) * sizeof(double));
        for (int i = 0; i <= n; ++i) { X[j][i] = j * config[10]; Y[j][i] = i * config[11]; }
    }
    double *cpu_time = calloc((size_t)N_T, sizeof(double));
    if (split)
        GRP_solver_2D_split_EUL_source(m, n, CV, X, Y, cpu_time, "glue", N_T, N_plot, time_plot);
    else
        GRP_solver_2D_EUL_source(m, n, CV, X, Y, cpu_time, "glue", N_T, N_plot, time_plot);
    for (int j = 0; j <= m; ++j) { free(X[j]); free(Y[j]); }
    free(X); free(Y); free(cpu_time); free(rows); free(CV);
}

/* Runs GRP_solver_2D_split_EUL_source. */
void ref_split2d(int m, int n, double *fields, int N_T, int *N_plot, double *time_plot)
{
    ref_2d(m, n, fields, N_T, N_plot, time_plot, 1);
}

/* Runs GRP_solver_2D_EUL_source (the unsplit step); arguments as ref_split2d. */
void ref_unsplit2d(int m, int n, double *fields, int N_T, int *N_plot, double *time_plot)
{
    ref_2d(m, n, fields, N_T, N_plot, time_plot, 0);
}

/* Runs GRP_solver_EUL_source (1D).  fields: 4*N_T blocks of m doubles (rho u p e). */
void ref_grp1d(int m, double *fields, int N_T, int *N_plot, double *time_plot)
{
    struct cell_var_stru CV;
    memset(&CV, 0, sizeof CV);
    double **rows = malloc((size_t)4 * N_T * sizeof(double *));
    for (int f = 0; f < 4; ++f)
        for (int k = 0; k < N_T; ++k) rows[f * N_T + k] = fields + ((size_t)f * N_T + k) * m;
    CV.RHO = rows; CV.U = rows + N_T; CV.P = rows + 2 * N_T; CV.E = rows + 3 * N_T;
    double *cpu_time = calloc((size_t)N_T, sizeof(double));
    *N_plot = N_T;
    GRP_solver_EUL_source(m, CV, cpu_time, N_plot, time_plot);
    free(cpu_time); free(rows);
}

/* ---- the 2D step's stages on their own: CV[0] (frame 0 + scratch), CV[1] (frame nt = 1) ---- */

static double **rows_of(double *block, int m, int n)
{
    double **r = malloc((size_t)m * sizeof(double *));
    for (int j = 0; j < m; ++j) r[j] = block + (size_t)j * n;
    return r;
}

/* fields f: 4*m*n doubles rho u v p -> frame k of CV */
static void set_frame(struct cell_var_stru *cv, double *f, int m, int n)
{
    const size_t mn = (size_t)m * n;
    cv->RHO = rows_of(f, m, n); cv->U = rows_of(f + mn, m, n);
    cv->V = rows_of(f + 2 * mn, m, n); cv->P = rows_of(f + 3 * mn, m, n);
}

static void free_frame(struct cell_var_stru *cv)
{
    free(cv->RHO); free(cv->U); free(cv->V); free(cv->P);
}

/* bound_cond_slope_limiter_{x,y} (dir 0/1) with nt = 1.  f0, f1: frames 0 and 1;
 * sl: the direction's slopes of CV[0] in (previous values) and out, 4*m*n;
 * cross: the other direction's slopes (untouched, 4*m*n); bfv*: b_f_var arrays
 * (13 doubles each) in and out.  config[] must be set. */
int ref_limiter(int dir, int m, int n, double *f0, double *f1, double *sl, double *cross, double *bfvL,
                double *bfvR, double *bfvD, double *bfvU, int found, int slope)
{
    struct cell_var_stru CV[2];
    memset(CV, 0, sizeof CV);
    set_frame(&CV[0], f0, m, n);
    set_frame(&CV[1], f1, m, n);
    const size_t mn = (size_t)m * n;
    double *sx = dir ? cross : sl, *sy = dir ? sl : cross;
    CV[0].s_rho = rows_of(sx, m, n); CV[0].s_u = rows_of(sx + mn, m, n);
    CV[0].s_v = rows_of(sx + 2 * mn, m, n); CV[0].s_p = rows_of(sx + 3 * mn, m, n);
    CV[0].t_rho = rows_of(sy, m, n); CV[0].t_u = rows_of(sy + mn, m, n);
    CV[0].t_v = rows_of(sy + 2 * mn, m, n); CV[0].t_p = rows_of(sy + 3 * mn, m, n);
    _Bool r = dir ? bound_cond_slope_limiter_y(m, n, 1, CV, (struct b_f_var *)bfvL, (struct b_f_var *)bfvR,
                                               (struct b_f_var *)bfvD, (struct b_f_var *)bfvU, found, slope, 0.0)
                  : bound_cond_slope_limiter_x(m, n, 1, CV, (struct b_f_var *)bfvL, (struct b_f_var *)bfvR,
                                               (struct b_f_var *)bfvD, (struct b_f_var *)bfvU, found, slope, 0.0);
    free(CV[0].s_rho); free(CV[0].s_u); free(CV[0].s_v); free(CV[0].s_p);
    free(CV[0].t_rho); free(CV[0].t_u); free(CV[0].t_v); free(CV[0].t_p);
    free_frame(&CV[0]);
    free_frame(&CV[1]);
    return r;
}

/* flux_generator_{x,y} (dir 0/1) with nt = 1 on frame f1, x-slopes s and
 * y-slopes t of CV[0]; bfv_lo / bfv_hi: b_f_var arrays (13 doubles each).
 * F, I: 4 * nif doubles (F_rho F_u F_v F_e / rho u v p at t + tau), interfaces
 * at [j][i] of (m+1) x n (x) or m x (n+1) (y).  Returns the function's value. */
int ref_flux_gen(int dir, int m, int n, double *f1, double *s, double *t, double *bfv_lo, double *bfv_hi,
                 double tau, int transversal, double *F, double *I)
{
    struct cell_var_stru CV[2];
    memset(CV, 0, sizeof CV);
    set_frame(&CV[1], f1, m, n);
    const size_t mn = (size_t)m * n;
    CV[0].s_rho = rows_of(s, m, n); CV[0].s_u = rows_of(s + mn, m, n);
    CV[0].s_v = rows_of(s + 2 * mn, m, n); CV[0].s_p = rows_of(s + 3 * mn, m, n);
    CV[0].t_rho = rows_of(t, m, n); CV[0].t_u = rows_of(t + mn, m, n);
    CV[0].t_v = rows_of(t + 2 * mn, m, n); CV[0].t_p = rows_of(t + 3 * mn, m, n);
    const int rows = dir ? m : m + 1, cols = dir ? n + 1 : n;
    const size_t nif = (size_t)rows * cols;
    double **fr[4], **ir[4];
    for (int q = 0; q < 4; ++q) {
        fr[q] = rows_of(F + q * nif, rows, cols);
        ir[q] = rows_of(I + q * nif, rows, cols);
    }
    int rc;
    if (dir) {
        CV[0].G_rho = fr[0]; CV[0].G_u = fr[1]; CV[0].G_v = fr[2]; CV[0].G_e = fr[3];
        CV[0].rhoIy = ir[0]; CV[0].uIy = ir[1]; CV[0].vIy = ir[2]; CV[0].pIy = ir[3];
        rc = flux_generator_y(m, n, 1, tau, CV, (struct b_f_var *)bfv_lo, (struct b_f_var *)bfv_hi, transversal);
    } else {
        CV[0].F_rho = fr[0]; CV[0].F_u = fr[1]; CV[0].F_v = fr[2]; CV[0].F_e = fr[3];
        CV[0].rhoIx = ir[0]; CV[0].uIx = ir[1]; CV[0].vIx = ir[2]; CV[0].pIx = ir[3];
        rc = flux_generator_x(m, n, 1, tau, CV, (struct b_f_var *)bfv_lo, (struct b_f_var *)bfv_hi, transversal);
    }
    for (int q = 0; q < 4; ++q) { free(fr[q]); free(ir[q]); }
    free(CV[0].s_rho); free(CV[0].s_u); free(CV[0].s_v); free(CV[0].s_p);
    free(CV[0].t_rho); free(CV[0].t_u); free(CV[0].t_v); free(CV[0].t_p);
    free_frame(&CV[1]);
    return rc;
}
