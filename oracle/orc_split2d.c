/*
 * orc_split2d.c -- the dimension-split 2D Eulerian GRP time loop, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_split2d        <- GRP_solver_2D_split_EUL_source  src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379
 *   bc_limiter_x/_y    <- bound_cond_slope_limiter_x/_y   src/inter_process/bound_cond_slope_limiter_x.c:27-129, _y.c:27-129
 *   limiter_line       <- minmod_limiter_2D_x / minmod_limiter (fixed h)
 *                                                          src/inter_process/slope_limiter_2D_x.c:32-84, slope_limiter.c:31-83
 *   flux_sweep_x/_y    <- flux_generator_x/_y (Transversal = false)
 *                                                          src/flux_calc/flux_generator_x.c:31-165, _y.c:31-165
 *
 * One deliberate deviation: when GRP_2D_flux fails its star/derivative check
 * the reference stores whatever flux the same OpenMP thread produced last (a
 * schedule-dependent value) and then stops after the update.  Here such an
 * interface stores zero flux and zero interface values; the stop step is the
 * same.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "grp_oracle.h"

/* ghost cell: struct b_f_var (src/include/var_struc.h:146-151) minus H */
typedef struct ghost {
    double RHO, P, U, V;
    double SRHO, SP, SU, SV;
    double TRHO, TP, TU, TV;
} ghost;

typedef struct grid2d {
    int m, n;
    double h_x, h_y, alpha, gamma, eps;
    int bc_x, bc_y;
    double *rho, *u, *v, *p, *e;               /* current frame        */
    const double *rho0, *u0, *v0, *p0;         /* frame 0 (initial)    */
    double *s[4], *t[4];                       /* x / y slopes: rho u v p */
    double *Fx[4], *Ix[4];                     /* (m+1)*n: [j*n + i]   */
    double *Gy[4], *Iy[4];                     /* m*(n+1): [j*(n+1)+i] */
    ghost *gL, *gR, *gD, *gU;
} grid2d;

#define AT(j, i) ((size_t)(j) * (size_t)g->n + (size_t)(i))

/* slope of one variable along a line of `cnt` cells; element k at base[k*stride] */
static void limiter_line(int cnt, size_t stride, const double *U, double *s, double UL, double UR,
                         double h, int have_prev, double alpha)
{
    for (int j = 0; j < cnt; ++j) {
        const double Uj = U[(size_t)j * stride];
        const double s_L = j ? (Uj - U[(size_t)(j - 1) * stride]) / h : (Uj - UL) / h;
        const double s_R = (j < cnt - 1) ? (U[(size_t)(j + 1) * stride] - Uj) / h : (UR - Uj) / h;
        double *sj = &s[(size_t)j * stride];
        *sj = have_prev ? orc_minmod3(alpha * s_L, alpha * s_R, *sj) : orc_minmod2(s_L, s_R);
    }
}

/* ghost values in x for line i (bound_cond_slope_limiter_x.c:34-82) */
static int ghost_values_x(grid2d *g, int i, int found)
{
    const int m = g->m;
    ghost *L = &g->gL[i], *R = &g->gR[i];
    switch (g->bc_x) {
    case -1:
        if (found) break;
        L->U = g->u0[AT(0, i)]; R->U = g->u0[AT(m - 1, i)];
        L->V = g->v0[AT(0, i)]; R->V = g->v0[AT(m - 1, i)];
        L->P = g->p0[AT(0, i)]; R->P = g->p0[AT(m - 1, i)];
        L->RHO = g->rho0[AT(0, i)]; R->RHO = g->rho0[AT(m - 1, i)];
        break;
    case -2: case -4: case -24: {
        const double sl = (g->bc_x == -4) ? 1.0 : -1.0;     /* left normal velocity  */
        const double sr = (g->bc_x == -2) ? -1.0 : 1.0;     /* right normal velocity */
        L->U = sl < 0 ? -g->u[AT(0, i)] : g->u[AT(0, i)];
        R->U = sr < 0 ? -g->u[AT(m - 1, i)] : g->u[AT(m - 1, i)];
        L->V = g->v[AT(0, i)]; R->V = g->v[AT(m - 1, i)];
        L->P = g->p[AT(0, i)]; R->P = g->p[AT(m - 1, i)];
        L->RHO = g->rho[AT(0, i)]; R->RHO = g->rho[AT(m - 1, i)];
        break;
    }
    case -7:
        L->U = g->u[AT(m - 1, i)]; R->U = g->u[AT(0, i)];
        L->V = g->v[AT(m - 1, i)]; R->V = g->v[AT(0, i)];
        L->P = g->p[AT(m - 1, i)]; R->P = g->p[AT(0, i)];
        L->RHO = g->rho[AT(m - 1, i)]; R->RHO = g->rho[AT(0, i)];
        break;
    default:
        return 0;
    }
    return 1;
}

/* ghost values in y for column j (bound_cond_slope_limiter_y.c:34-82) */
static int ghost_values_y(grid2d *g, int j, int found)
{
    const int n = g->n;
    ghost *D = &g->gD[j], *U = &g->gU[j];
    switch (g->bc_y) {
    case -1:
        if (found) break;
        D->U = g->u0[AT(j, 0)]; U->U = g->u0[AT(j, n - 1)];
        D->V = g->v0[AT(j, 0)]; U->V = g->v0[AT(j, n - 1)];
        D->P = g->p0[AT(j, 0)]; U->P = g->p0[AT(j, n - 1)];
        D->RHO = g->rho0[AT(j, 0)]; U->RHO = g->rho0[AT(j, n - 1)];
        break;
    case -2: case -4: case -24: {
        const double sd = (g->bc_y == -4) ? 1.0 : -1.0;
        const double su = (g->bc_y == -2) ? -1.0 : 1.0;
        D->U = g->u[AT(j, 0)]; U->U = g->u[AT(j, n - 1)];
        D->V = sd < 0 ? -g->v[AT(j, 0)] : g->v[AT(j, 0)];
        U->V = su < 0 ? -g->v[AT(j, n - 1)] : g->v[AT(j, n - 1)];
        D->P = g->p[AT(j, 0)]; U->P = g->p[AT(j, n - 1)];
        D->RHO = g->rho[AT(j, 0)]; U->RHO = g->rho[AT(j, n - 1)];
        break;
    }
    case -7:
        D->U = g->u[AT(j, n - 1)]; U->U = g->u[AT(j, 0)];
        D->V = g->v[AT(j, n - 1)]; U->V = g->v[AT(j, 0)];
        D->P = g->p[AT(j, n - 1)]; U->P = g->p[AT(j, 0)];
        D->RHO = g->rho[AT(j, n - 1)]; U->RHO = g->rho[AT(j, 0)];
        break;
    default:
        return 0;
    }
    return 1;
}

/* bound_cond_slope_limiter_x: ghosts, limited x-slopes, ghost slopes */
static int bc_limiter_x(grid2d *g, int found)
{
    const int m = g->m, n = g->n;
    for (int i = 0; i < n; ++i)
        if (!ghost_values_x(g, i, found)) {
            printf("No suitable boundary coditions in x direction!\n");
            return 0;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const ghost *L = &g->gL[i], *R = &g->gR[i];
        limiter_line(m, (size_t)n, g->u + i,   g->s[1] + i, L->U,   R->U,   g->h_x, found, g->alpha);
        limiter_line(m, (size_t)n, g->v + i,   g->s[2] + i, L->V,   R->V,   g->h_x, found, g->alpha);
        limiter_line(m, (size_t)n, g->p + i,   g->s[3] + i, L->P,   R->P,   g->h_x, found, g->alpha);
        limiter_line(m, (size_t)n, g->rho + i, g->s[0] + i, L->RHO, R->RHO, g->h_x, found, g->alpha);
    }
    for (int i = 0; i < n; ++i) {
        ghost *L = &g->gL[i], *R = &g->gR[i];
        switch (g->bc_x) {
        case -2:
            L->SU = g->s[1][AT(0, i)]; R->SU = g->s[1][AT(m - 1, i)];
            break;
        case -7:
            L->SU = g->s[1][AT(m - 1, i)];   R->SU = g->s[1][AT(0, i)];
            L->SV = g->s[2][AT(m - 1, i)];   R->SV = g->s[2][AT(0, i)];
            L->SP = g->s[3][AT(m - 1, i)];   R->SP = g->s[3][AT(0, i)];
            L->SRHO = g->s[0][AT(m - 1, i)]; R->SRHO = g->s[0][AT(0, i)];
            break;
        case -24:
            L->SU = g->s[1][AT(0, i)];
            break;
        }
    }
    /* x-slopes at the y-boundaries feed only the transversal (unsplit) terms */
    return 1;
}

static int bc_limiter_y(grid2d *g, int found)
{
    const int m = g->m, n = g->n;
    for (int j = 0; j < m; ++j)
        if (!ghost_values_y(g, j, found)) {
            printf("No suitable boundary coditions in y direction!\n");
            return 0;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < m; ++j) {
        const ghost *D = &g->gD[j], *U = &g->gU[j];
        const size_t o = AT(j, 0);
        limiter_line(n, 1, g->u + o,   g->t[1] + o, D->U,   U->U,   g->h_y, found, g->alpha);
        limiter_line(n, 1, g->v + o,   g->t[2] + o, D->V,   U->V,   g->h_y, found, g->alpha);
        limiter_line(n, 1, g->p + o,   g->t[3] + o, D->P,   U->P,   g->h_y, found, g->alpha);
        limiter_line(n, 1, g->rho + o, g->t[0] + o, D->RHO, U->RHO, g->h_y, found, g->alpha);
    }
    for (int j = 0; j < m; ++j) {
        ghost *D = &g->gD[j], *U = &g->gU[j];
        switch (g->bc_y) {
        case -2:
            D->TV = g->t[2][AT(j, 0)]; U->TV = g->t[2][AT(j, n - 1)];
            break;
        case -7:
            D->TU = g->t[1][AT(j, n - 1)];   U->TU = g->t[1][AT(j, 0)];
            D->TV = g->t[2][AT(j, n - 1)];   U->TV = g->t[2][AT(j, 0)];
            D->TP = g->t[3][AT(j, n - 1)];   U->TP = g->t[3][AT(j, 0)];
            D->TRHO = g->t[0][AT(j, n - 1)]; U->TRHO = g->t[0][AT(j, 0)];
            break;
        case -24:
            D->TV = g->t[2][AT(j, 0)];
            break;
        }
    }
    return 1;
}

/* interface state from a cell (or ghost) value and slope: value +- 0.5*h*slope */
static void side_state(orc_ifv *f, double gamma, double rho, double u, double v, double p,
                       double sr, double su, double sv, double sp, double half_h_sign, double tz)
{
    f->gamma = gamma; f->lambda_u = 0.0; f->lambda_v = 0.0;
    f->d_rho = sr; f->d_u = su; f->d_v = sv; f->d_p = sp;
    if (half_h_sign > 0) {
        f->rho = rho + half_h_sign * sr; f->u = u + half_h_sign * su;
        f->v = v + half_h_sign * sv;     f->p = p + half_h_sign * sp;
    } else {
        f->rho = rho - (-half_h_sign) * sr; f->u = u - (-half_h_sign) * su;
        f->v = v - (-half_h_sign) * sv;     f->p = p - (-half_h_sign) * sp;
    }
    f->t_rho = tz; f->t_u = tz; f->t_v = tz; f->t_p = tz;
}

/* flux_generator_x with Transversal = false; returns 0, 1 (reconstruction) or 2 (GRP) */
static int flux_sweep_x(grid2d *g, double tau)
{
    const int m = g->m, n = g->n;
    const double hh = 0.5 * g->h_x;
    long long last = -1;   /* (loop-order index << 2) | code of the last failing interface */
#pragma omp parallel for collapse(2) schedule(static) reduction(max:last)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= m; ++j) {
            int code = 0;
            orc_ifv L, R;
            if (j) {
                const size_t c = AT(j - 1, i);
                side_state(&L, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                           g->s[0][c], g->s[1][c], g->s[2][c], g->s[3][c], hh, 0.0);
            } else {
                const ghost *b = &g->gL[i];
                side_state(&L, g->gamma, b->RHO, b->U, b->V, b->P, b->SRHO, b->SU, b->SV, b->SP, hh, 0.0);
            }
            if (j < m) {
                const size_t c = AT(j, i);
                side_state(&R, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                           g->s[0][c], g->s[1][c], g->s[2][c], g->s[3][c], -hh, 0.0);
            } else {
                const ghost *b = &g->gR[i];
                side_state(&R, g->gamma, b->RHO, b->U, b->V, b->P, b->SRHO, b->SU, b->SV, b->SP, -hh, 0.0);
            }
            if (orc_ifvar_check2(&L, &R, g->eps)) {
                printf("<0.0 or non-finite reconstruction on [x, %d, %d]\n", j, i);
                code = 1;
            }
            double F[4] = {0, 0, 0, 0}, I[4] = {0, 0, 0, 0};
            if (orc_grp_2d_flux(L, R, tau, g->eps, 1.0, 0.0, F, I)) {
                printf("GRP star/derivative error on [x, %d, %d]\n", j, i);
                code = 2;
            }
            if (code) {
                const long long key = (((long long)i * (m + 1) + j) << 2) | code;
                if (key > last) last = key;
            }
            const size_t f = (size_t)j * n + i;
            for (int q = 0; q < 4; ++q) { g->Fx[q][f] = F[q]; g->Ix[q][f] = I[q]; }
        }
    return last < 0 ? 0 : (int)(last & 3);
}

/* flux_generator_y with Transversal = false (t_* = -0.0) */
static int flux_sweep_y(grid2d *g, double tau)
{
    const int m = g->m, n = g->n;
    const double hh = 0.5 * g->h_y;
    long long last = -1;
#pragma omp parallel for collapse(2) schedule(static) reduction(max:last)
    for (int j = 0; j < m; ++j)
        for (int i = 0; i <= n; ++i) {
            int code = 0;
            orc_ifv D, U;
            if (i) {
                const size_t c = AT(j, i - 1);
                side_state(&D, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                           g->t[0][c], g->t[1][c], g->t[2][c], g->t[3][c], hh, -0.0);
            } else {
                const ghost *b = &g->gD[j];
                side_state(&D, g->gamma, b->RHO, b->U, b->V, b->P, b->TRHO, b->TU, b->TV, b->TP, hh, -0.0);
            }
            if (i < n) {
                const size_t c = AT(j, i);
                side_state(&U, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                           g->t[0][c], g->t[1][c], g->t[2][c], g->t[3][c], -hh, -0.0);
            } else {
                const ghost *b = &g->gU[j];
                side_state(&U, g->gamma, b->RHO, b->U, b->V, b->P, b->TRHO, b->TU, b->TV, b->TP, -hh, -0.0);
            }
            if (orc_ifvar_check2(&D, &U, g->eps)) {
                printf("<0.0 or non-finite reconstruction on [y, %d, %d]\n", j, i);
                code = 1;
            }
            double F[4] = {0, 0, 0, 0}, I[4] = {0, 0, 0, 0};
            if (orc_grp_2d_flux(D, U, tau, g->eps, 0.0, 1.0, F, I)) {
                printf("GRP star/derivative error on [y, %d, %d]\n", j, i);
                code = 2;
            }
            if (code) {
                const long long key = (((long long)j * (n + 1) + i) << 2) | code;
                if (key > last) last = key;
            }
            const size_t f = (size_t)j * (n + 1) + i;
            for (int q = 0; q < 4; ++q) { g->Gy[q][f] = F[q]; g->Iy[q][f] = I[q]; }
        }
    return last < 0 ? 0 : (int)(last & 3);
}

/* conservative update + x-slope refresh (:229-255); returns 1 if positivity failed */
static int update_x(grid2d *g, double half_nu)
{
    const int m = g->m, n = g->n;
    const double gm1 = g->gamma - 1.0;
    int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(|:bad)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const size_t c = AT(j, i), f0 = (size_t)j * n + i, f1 = (size_t)(j + 1) * n + i;
            const double mom_x = g->rho[c] * g->u[c] - half_nu * (g->Fx[1][f1] - g->Fx[1][f0]);
            const double mom_y = g->rho[c] * g->v[c] - half_nu * (g->Fx[2][f1] - g->Fx[2][f0]);
            const double ene   = g->rho[c] * g->e[c] - half_nu * (g->Fx[3][f1] - g->Fx[3][f0]);
            g->rho[c] += -half_nu * (g->Fx[0][f1] - g->Fx[0][f0]);
            g->u[c] = mom_x / g->rho[c];
            g->v[c] = mom_y / g->rho[c];
            g->e[c] = ene / g->rho[c];
            g->p[c] = (ene - 0.5 * mom_x * g->u[c] - 0.5 * mom_y * g->v[c]) * gm1;
            if (g->p[c] < g->eps || g->rho[c] < g->eps)
                bad = 1;
            for (int q = 0; q < 4; ++q)
                g->s[q][c] = (g->Ix[q][f1] - g->Ix[q][f0]) / g->h_x;
        }
    return bad;
}

static int update_y(grid2d *g, double mu)
{
    const int m = g->m, n = g->n;
    const double gm1 = g->gamma - 1.0;
    int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(|:bad)
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
            const size_t c = AT(j, i), f0 = (size_t)j * (n + 1) + i, f1 = f0 + 1;
            const double mom_x = g->rho[c] * g->u[c] - mu * (g->Gy[1][f1] - g->Gy[1][f0]);
            const double mom_y = g->rho[c] * g->v[c] - mu * (g->Gy[2][f1] - g->Gy[2][f0]);
            const double ene   = g->rho[c] * g->e[c] - mu * (g->Gy[3][f1] - g->Gy[3][f0]);
            g->rho[c] += -mu * (g->Gy[0][f1] - g->Gy[0][f0]);
            g->u[c] = mom_x / g->rho[c];
            g->v[c] = mom_y / g->rho[c];
            g->e[c] = ene / g->rho[c];
            g->p[c] = (ene - 0.5 * mom_x * g->u[c] - 0.5 * mom_y * g->v[c]) * gm1;
            if (g->p[c] < g->eps || g->rho[c] < g->eps)
                bad = 1;
            for (int q = 0; q < 4; ++q)
                g->t[q][c] = (g->Iy[q][f1] - g->Iy[q][f0]) / g->h_y;
        }
    return bad;
}

static double wall_time(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / (double)CLOCKS_PER_SEC;
#endif
}

int orc_split2d(int m, int n, double *const *rho, double *const *u, double *const *v,
                double *const *p, double *const *e, int N_T, int *N_plot,
                double *time_plot, double *cpu_time, double *config)
{
    const double t_all = config[1], eps = config[4], gamma = config[6], CFL = config[7];
    const int N = (int)config[5];
    const double h_x = config[10], h_y = config[11];
    double tau = config[16];

    grid2d G, *g = &G;
    memset(g, 0, sizeof G);
    g->m = m; g->n = n; g->h_x = h_x; g->h_y = h_y; g->alpha = config[41];
    g->gamma = gamma; g->eps = eps;
    g->bc_x = (int)config[17]; g->bc_y = (int)config[18];
    g->rho0 = rho[0]; g->u0 = u[0]; g->v0 = v[0]; g->p0 = p[0];
    const size_t mn = (size_t)m * n, fx = (size_t)(m + 1) * n, fy = (size_t)m * (n + 1);
    int status = 0;
    for (int q = 0; q < 4; ++q) {
        g->s[q] = calloc(mn, sizeof(double)); g->t[q] = calloc(mn, sizeof(double));
        g->Fx[q] = calloc(fx, sizeof(double)); g->Ix[q] = calloc(fx, sizeof(double));
        g->Gy[q] = calloc(fy, sizeof(double)); g->Iy[q] = calloc(fy, sizeof(double));
        if (!g->s[q] || !g->t[q] || !g->Fx[q] || !g->Ix[q] || !g->Gy[q] || !g->Iy[q]) status = 1;
    }
    g->gL = calloc((size_t)n, sizeof(ghost)); g->gR = calloc((size_t)n, sizeof(ghost));
    g->gD = calloc((size_t)m, sizeof(ghost)); g->gU = calloc((size_t)m, sizeof(ghost));
    if (!g->gL || !g->gR || !g->gD || !g->gU) status = 1;

    int k = 0, nt = 0, nt_plot = 0, DS = 1, found_x = 0, found_y = 0, stop_t = 0;
    double time_c = 0.0, half_tau = 0.0, half_nu = 0.0, mu = 0.0, cpu_sum = 0.0;
    if (status) { printf("NOT enough memory!\n"); goto done; }

    for (k = 1; k <= N; DS ? k : ++k) {
        const double tic = wall_time();
        if (time_c >= time_plot[nt_plot] && nt_plot < (*N_plot - 1)) {      /* :157-176 */
            nt_plot++;
            if (nt < N_T - 1) {
                memcpy(rho[nt + 1], rho[nt], mn * sizeof(double));
                memcpy(u[nt + 1], u[nt], mn * sizeof(double));
                memcpy(v[nt + 1], v[nt], mn * sizeof(double));
                memcpy(e[nt + 1], e[nt], mn * sizeof(double));
                memcpy(p[nt + 1], p[nt], mn * sizeof(double));
                nt++;
            }
        }
        g->rho = rho[nt]; g->u = u[nt]; g->v = v[nt]; g->p = p[nt]; g->e = e[nt];

        if (DS) {                                                          /* :182-212 */
            double h_S_max = INFINITY;
            for (size_t c = 0; c < mn; ++c) {
                const double cs = sqrt(gamma * g->p0[c] / g->rho0[c]);
                const double sigma = fabs(cs) + fabs(g->u0[c]) + fabs(g->v0[c]);
                h_S_max = fmin(h_S_max, fmin(h_x, h_y) / sigma);
            }
            if (isfinite(t_all) || !isfinite(config[16]) || config[16] <= 0.0) {
                tau = CFL * h_S_max;
                if (tau < eps) {
                    printf("\nThe length of the time step is so small on [%d, %g, %g] (t_n, time_c, tau)\n", k, time_c, tau);
                    stop_t = 1;
                } else if ((time_c + tau) > (t_all - eps)) {
                    tau = t_all - time_c;
                } else if (!isfinite(tau)) {
                    printf("NAN or INFinite error on [%d, %g, %g] (t_n, time_c, tau) - CFL\n", k, time_c, tau);
                    status = 1;
                    goto done;
                }
            }
            half_tau = tau * 0.5;
            half_nu = half_tau / h_x;
            mu = tau / h_y;
        }

        if (!bc_limiter_x(g, found_x)) { status = 1; goto done; }
        found_x = 1;
        int ferr = flux_sweep_x(g, half_tau);
        if (ferr == 1) { status = 1; goto done; }
        if (ferr == 2) stop_t = 1;
        if (update_x(g, half_nu)) { printf("<0.0 error on step %d - Update x\n", k); stop_t = 1; }
        if (stop_t) break;

        if (DS) {
            if (!bc_limiter_y(g, found_y)) { status = 1; goto done; }
            found_y = 1;
            ferr = flux_sweep_y(g, tau);
            if (ferr == 1) { status = 1; goto done; }
            if (ferr == 2) stop_t = 1;
            if (update_y(g, mu)) { printf("<0.0 error on step %d - Update y\n", k); stop_t = 1; }
            time_c += tau;
        }
        if (stop_t || time_c > (t_all - eps) || !isfinite(time_c))
            break;
        DS = DS ? 0 : 1;
        cpu_sum += wall_time() - tic;
        if (cpu_time) cpu_time[nt] = cpu_sum;
    }
    if (stop_t) status = 2;

done:
    config[5] = (double)k;
    *N_plot = nt + 1;
    if (isfinite(time_c))      time_plot[nt] = time_c;
    else if (isfinite(t_all))  time_plot[nt] = t_all;
    else if (isfinite(tau))    time_plot[nt] = k * tau;
    for (int q = 0; q < 4; ++q) {
        free(g->s[q]); free(g->t[q]); free(g->Fx[q]); free(g->Ix[q]); free(g->Gy[q]); free(g->Iy[q]);
    }
    free(g->gL); free(g->gR); free(g->gD); free(g->gU);
    return status;
}

/* threads the OpenMP loops of this library will use (for reporting) */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
