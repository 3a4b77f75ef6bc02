/*
 * orc_grid2d.h -- grid, ghost cells, slope limiters and interface
 * reconstruction shared by the two 2D time loops of the oracle (split and
 * unsplit).  TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   bc_limiter_x/_y    <- bound_cond_slope_limiter_x/_y   src/inter_process/bound_cond_slope_limiter_x.c:27-129, _y.c:27-129
 *   limiter_line       <- minmod_limiter_2D_x / minmod_limiter (fixed h)
 *                                                          src/inter_process/slope_limiter_2D_x.c:32-84, slope_limiter.c:31-83
 */
#ifndef ORC_GRID2D_H
#define ORC_GRID2D_H
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
    /* x-slopes of the cells next to the y-boundaries: the transversal
     * derivatives of the y-ghosts (bound_cond_slope_limiter_x.c:111-126); read
     * by the unsplit step only */
    for (int j = 0; j < m; ++j) {
        ghost *D = &g->gD[j], *U = &g->gU[j];
        int lo, hi;
        if (g->bc_y == -2 || g->bc_y == -4 || g->bc_y == -24) { lo = 0; hi = n - 1; }
        else if (g->bc_y == -7) { lo = n - 1; hi = 0; }
        else continue;
        D->SU = g->s[1][AT(j, lo)];   U->SU = g->s[1][AT(j, hi)];
        D->SV = g->s[2][AT(j, lo)];   U->SV = g->s[2][AT(j, hi)];
        D->SP = g->s[3][AT(j, lo)];   U->SP = g->s[3][AT(j, hi)];
        D->SRHO = g->s[0][AT(j, lo)]; U->SRHO = g->s[0][AT(j, hi)];
    }
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
    /* y-slopes of the cells next to the x-boundaries: the transversal
     * derivatives of the x-ghosts (bound_cond_slope_limiter_y.c:111-126) */
    for (int i = 0; i < n; ++i) {
        ghost *L = &g->gL[i], *R = &g->gR[i];
        int lo, hi;
        if (g->bc_x == -2 || g->bc_x == -4 || g->bc_x == -24) { lo = 0; hi = m - 1; }
        else if (g->bc_x == -7) { lo = m - 1; hi = 0; }
        else continue;
        L->TU = g->t[1][AT(lo, i)];   R->TU = g->t[1][AT(hi, i)];
        L->TV = g->t[2][AT(lo, i)];   R->TV = g->t[2][AT(hi, i)];
        L->TP = g->t[3][AT(lo, i)];   R->TP = g->t[3][AT(hi, i)];
        L->TRHO = g->t[0][AT(lo, i)]; R->TRHO = g->t[0][AT(hi, i)];
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

static double wall_time(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / (double)CLOCKS_PER_SEC;
#endif
}

#endif
