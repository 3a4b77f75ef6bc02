/*
 * lib_stages.c -- the 2D step's stages with the reference's names and
 * signatures: bound_cond_slope_limiter_x / _y (src/include/inter_process.h:51,
 * 56) and flux_generator_x / _y (src/include/flux_calc.h:16, 21).
 *
 * They read and write the same CV members as the reference's (CV[nt] state,
 * CV[0]'s slope, flux and interface arrays, the b_f_var ghost arrays).  Ghost
 * values and ghost slopes are a few assignments per line and stay on the
 * host; the slopes over the grid and the interface solves run on the GPU
 * (hgrp2d_slopes, hgrp2d_fluxes in include/hydrocode_grp.h), so each call pays
 * a host <-> device copy of the grid: these are for callers that use the
 * stages one by one; the time step itself fuses them on the device.
 *
 * Deliberate difference: where an interface's solver check fails, the
 * reference stores whatever the previous interface of the same OpenMP thread
 * left in its i_f_var (schedule-dependent); these store NaN.  The return
 * value is the reference's: the code of the last failing interface in its
 * loop order (1: ifvar_check, 2: the solver's check).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrocode_dropin.h"
#include "hydrocode_grp.h"

static const char *bc_text(int bc)
{
    switch (bc) {
    case -2: return "Reflective";
    case -4: return "Free";
    case -7: return "Periodic";
    case -24: return "Reflective + Free";
    default: return NULL;
    }
}

/* ghost b of a line from the cell values (rho, u, v, p); sgn_u / sgn_v negate a velocity */
static void set_ghost(struct b_f_var *b, double rho, double u, double v, double p, double sgn_u, double sgn_v)
{
    b->RHO = rho;
    b->U = sgn_u * u;
    b->V = sgn_v * v;
    b->P = p;
}

/* The ghost values of one direction (bound_cond_slope_limiter_x.c:34-82, _y.c:34-82).
 * dir 0: lines i < n along x, ghosts lo = bfv_L, hi = bfv_R; dir 1: lines j < m, lo = bfv_D, hi = bfv_U.
 * Returns 0 for an unknown boundary condition. */
static int ghost_values(int dir, int m, int n, int nt, const struct cell_var_stru *CV, struct b_f_var *lo,
                        struct b_f_var *hi, int found, double t_c)
{
    const int bc = (int)config[17 + dir];
    const char axis = dir ? 'y' : 'x';
    const int lines = dir ? m : n, len = dir ? n : m;
    if (bc == -1) {
        if (found)
            return 1;
        printf("Initial boudary conditions in %c direction at time %g .\n", axis, t_c);
    } else if (bc_text(bc)) {
        if (!found)
            printf("%s boudary conditions in %c direction.\n", bc_text(bc), axis);
    } else {
        printf("No suitable boundary coditions in %c direction!\n", axis);
        return 0;
    }
    /* -1 reads CV[0]'s initial data, the others the current frame */
    const struct cell_var_stru *F = bc == -1 ? CV : CV + nt;
    for (int l = 0; l < lines; ++l) {
        /* source cells of the two ghosts: own end, or the far end for periodic */
        const int a = bc == -7 ? len - 1 : 0, b = bc == -7 ? 0 : len - 1;
        const int ja = dir ? l : a, ia = dir ? a : l, jb = dir ? l : b, ib = dir ? b : l;
        /* reflective: the normal velocity changes sign (both ends; -24 the low end only) */
        const double flip_lo = (bc == -2 || bc == -24) ? -1.0 : 1.0, flip_hi = bc == -2 ? -1.0 : 1.0;
        set_ghost(lo + l, F->RHO[ja][ia], F->U[ja][ia], F->V[ja][ia], F->P[ja][ia], dir ? 1.0 : flip_lo,
                  dir ? flip_lo : 1.0);
        set_ghost(hi + l, F->RHO[jb][ib], F->U[jb][ib], F->V[jb][ib], F->P[jb][ib], dir ? 1.0 : flip_hi,
                  dir ? flip_hi : 1.0);
    }
    return 1;
}

static double *pack(int m, int n, double **const f[4])
{
    const size_t mn = (size_t)m * n;
    double *a = (double *)malloc(4 * mn * sizeof(double));
    if (!a)
        return NULL;
    for (int q = 0; q < 4; ++q)
        for (int j = 0; j < m; ++j)
            memcpy(a + q * mn + (size_t)j * n, f[q][j], (size_t)n * sizeof(double));
    return a;
}

static void unpack(int m, int n, const double *a, double **const f[4])
{
    const size_t mn = (size_t)m * n;
    for (int q = 0; q < 4; ++q)
        for (int j = 0; j < m; ++j)
            memcpy(f[q][j], a + q * mn + (size_t)j * n, (size_t)n * sizeof(double));
}

/* s (dir 0) or t (dir 1) of a ghost, rho u v p */
static double *gslope(struct b_f_var *b, int dir, int q)
{
    double *s[4] = {&b->SRHO, &b->SU, &b->SV, &b->SP}, *t[4] = {&b->TRHO, &b->TU, &b->TV, &b->TP};
    return dir ? t[q] : s[q];
}

static _Bool limiter(int dir, const int m, const int n, const int nt, struct cell_var_stru *CV,
                     struct b_f_var *bfv_L, struct b_f_var *bfv_R, struct b_f_var *bfv_D, struct b_f_var *bfv_U,
                     _Bool found, const _Bool Slope, const double t_c)
{
    struct b_f_var *lo = dir ? bfv_D : bfv_L, *hi = dir ? bfv_U : bfv_R;
    if (!ghost_values(dir, m, n, nt, CV, lo, hi, found, t_c))
        return 0;
    if (!Slope)
        return 1;
    const int lines = dir ? m : n, len = dir ? n : m;
    double **const st[4] = {CV[nt].RHO, CV[nt].U, CV[nt].V, CV[nt].P};
    double **const sl[4] = {dir ? CV->t_rho : CV->s_rho, dir ? CV->t_u : CV->s_u, dir ? CV->t_v : CV->s_v,
                            dir ? CV->t_p : CV->s_p};
    const size_t mn = (size_t)m * n;
    double *state = pack(m, n, st), *prev = found ? pack(m, n, sl) : NULL;
    double *out = (double *)malloc(4 * mn * sizeof(double));
    double *g = (double *)malloc((size_t)8 * lines * sizeof(double));
    int rc = (!state || !out || !g || (found && !prev)) ? -2 : 0;
    if (!rc) {
        for (int l = 0; l < lines; ++l) {
            const double vl[4] = {lo[l].RHO, lo[l].U, lo[l].V, lo[l].P}, vh[4] = {hi[l].RHO, hi[l].U, hi[l].V, hi[l].P};
            memcpy(g + 4 * l, vl, sizeof vl);
            memcpy(g + 4 * (lines + l), vh, sizeof vh);
        }
        rc = hgrp2d_slopes(dir, m, n, state, g, g + 4 * lines, prev, found, config[41], config[10 + dir], out);
    }
    if (!rc)
        unpack(m, n, out, sl);
    free(state);
    free(prev);
    free(out);
    free(g);
    if (rc) {
        printf("GPU slope limiter failed (%d)\n", rc);
        return 0;
    }
    /* ghost slopes along dir (bound_cond_slope_limiter_x.c:94-109, _y.c:94-109):
     * -2 the normal velocity's only, -7 all from the far end, -24 the low end's normal velocity */
    const int bc = (int)config[17 + dir], un = dir ? 2 : 1;
    for (int l = 0; l < lines; ++l) {
        const int ja = dir ? l : 0, ia = dir ? 0 : l, jb = dir ? l : len - 1, ib = dir ? len - 1 : l;
        for (int q = 0; q < 4; ++q) {
            if (bc == -7) {
                *gslope(lo + l, dir, q) = sl[q][jb][ib];
                *gslope(hi + l, dir, q) = sl[q][ja][ia];
            } else if ((bc == -2 || bc == -24) && q == un) {
                *gslope(lo + l, dir, q) = sl[q][ja][ia];
                if (bc == -2)
                    *gslope(hi + l, dir, q) = sl[q][jb][ib];
            }
        }
    }
    /* the other direction's ghosts get these slopes too, in the same members
     * (bound_cond_slope_limiter_x.c:111-126, _y.c:111-126), for the transversal
     * derivatives at its ends: free / reflective copy the edge cell's, periodic the far edge's */
    const int obc = (int)config[17 + !dir], olines = dir ? n : m, olen = dir ? m : n;
    struct b_f_var *olo = dir ? bfv_L : bfv_D, *ohi = dir ? bfv_R : bfv_U;
    if (obc == -2 || obc == -4 || obc == -24 || obc == -7) {
        for (int l = 0; l < olines; ++l) {
            const int a = obc == -7 ? olen - 1 : 0, b = obc == -7 ? 0 : olen - 1;
            const int ja = dir ? a : l, ia = dir ? l : a, jb = dir ? b : l, ib = dir ? l : b;
            for (int q = 0; q < 4; ++q) {
                *gslope(olo + l, dir, q) = sl[q][ja][ia];
                *gslope(ohi + l, dir, q) = sl[q][jb][ib];
            }
        }
    }
    return 1;
}

_Bool bound_cond_slope_limiter_x(const int m, const int n, const int nt, struct cell_var_stru *CV,
                                 struct b_f_var *bfv_L, struct b_f_var *bfv_R, struct b_f_var *bfv_D,
                                 struct b_f_var *bfv_U, _Bool find_bound_x, const _Bool Slope, const double t_c)
{
    return limiter(0, m, n, nt, CV, bfv_L, bfv_R, bfv_D, bfv_U, find_bound_x, Slope, t_c);
}

_Bool bound_cond_slope_limiter_y(const int m, const int n, const int nt, struct cell_var_stru *CV,
                                 struct b_f_var *bfv_L, struct b_f_var *bfv_R, struct b_f_var *bfv_D,
                                 struct b_f_var *bfv_U, _Bool find_bound_y, const _Bool Slope, const double t_c)
{
    return limiter(1, m, n, nt, CV, bfv_L, bfv_R, bfv_D, bfv_U, find_bound_y, Slope, t_c);
}

static int fluxes(int dir, const int m, const int n, const int nt, const double tau, struct cell_var_stru *CV,
                  struct b_f_var *lo, struct b_f_var *hi, const _Bool Transversal)
{
    const int lines = dir ? m : n;
    const size_t nif = dir ? (size_t)m * (n + 1) : (size_t)(m + 1) * n;
    double **const st[4] = {CV[nt].RHO, CV[nt].U, CV[nt].V, CV[nt].P};
    double **const s[4] = {CV->s_rho, CV->s_u, CV->s_v, CV->s_p};
    double **const t[4] = {CV->t_rho, CV->t_u, CV->t_v, CV->t_p};
    double *state = pack(m, n, st), *sn = pack(m, n, dir ? t : s), *tr = Transversal ? pack(m, n, dir ? s : t) : NULL;
    double *b = (double *)malloc((size_t)24 * lines * sizeof(double));
    double *F = (double *)malloc(4 * nif * sizeof(double)), *I = (double *)malloc(4 * nif * sizeof(double));
    int *codes = (int *)malloc(nif * sizeof(int));
    int rc = (!state || !sn || (Transversal && !tr) || !b || !F || !I || !codes) ? -2 : 0;
    if (!rc) {
        for (int l = 0; l < 2 * lines; ++l) {
            const struct b_f_var *g = l < lines ? lo + l : hi + (l - lines);
            const double v[12] = {g->RHO, g->U, g->V, g->P,
                                  dir ? g->TRHO : g->SRHO, dir ? g->TU : g->SU, dir ? g->TV : g->SV, dir ? g->TP : g->SP,
                                  dir ? g->SRHO : g->TRHO, dir ? g->SU : g->TU, dir ? g->SV : g->TV, dir ? g->SP : g->TP};
            memcpy(b + 12 * l, v, sizeof v);
        }
        rc = hgrp2d_fluxes(dir, m, n, state, sn, tr, b, b + 12 * lines, tau, config[10 + dir], config[6], config[4],
                           Transversal, F, I, codes);
    }
    int retval = 0;
    if (!rc) {
        double **const fo[4] = {dir ? CV->G_rho : CV->F_rho, dir ? CV->G_u : CV->F_u, dir ? CV->G_v : CV->F_v,
                                dir ? CV->G_e : CV->F_e};
        double **const io[4] = {dir ? CV->rhoIy : CV->rhoIx, dir ? CV->uIy : CV->uIx, dir ? CV->vIy : CV->vIx,
                                dir ? CV->pIy : CV->pIx};
        const int rows = dir ? m : m + 1, cols = dir ? n + 1 : n;
        for (int j = 0; j < rows; ++j)
            for (int q = 0; q < 4; ++q) {
                memcpy(fo[q][j], F + q * nif + (size_t)j * cols, (size_t)cols * sizeof(double));
                memcpy(io[q][j], I + q * nif + (size_t)j * cols, (size_t)cols * sizeof(double));
            }
        /* messages and return value in the reference's loop order (x: i outer, j inner; y: j outer, i inner) */
        const char d = dir ? 'y' : 'x';
        for (int outer = 0; outer < (dir ? m : n); ++outer)
            for (int inner = 0; inner < (dir ? n + 1 : m + 1); ++inner) {
                const int j = dir ? outer : inner, i = dir ? inner : outer;
                const int c = codes[(size_t)j * cols + i];
                if (c == 1)
                    printf("Interface state error on [%d, %d, %d] (nt, x, y).\n", nt, j, i);
                else if (c == 2)
                    printf("NAN or INFinite error on [%d, %d, %d] (nt, x, y) - STA_%c\n", nt, j, i, d);
                if (c)
                    retval = c;
            }
    } else {
        printf("GPU flux generator failed (%d)\n", rc);
        retval = 2;
    }
    free(state);
    free(sn);
    free(tr);
    free(b);
    free(F);
    free(I);
    free(codes);
    return retval;
}

int flux_generator_x(const int m, const int n, const int nt, const double tau, struct cell_var_stru *CV,
                     struct b_f_var *bfv_L, struct b_f_var *bfv_R, const _Bool Transversal)
{
    return fluxes(0, m, n, nt, tau, CV, bfv_L, bfv_R, Transversal);
}

int flux_generator_y(const int m, const int n, const int nt, const double tau, struct cell_var_stru *CV,
                     struct b_f_var *bfv_D, struct b_f_var *bfv_U, const _Bool Transversal)
{
    return fluxes(1, m, n, nt, tau, CV, bfv_D, bfv_U, Transversal);
}
