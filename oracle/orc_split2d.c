/*
 * orc_split2d.c -- the dimension-split 2D Eulerian GRP time loop, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_split2d        <- GRP_solver_2D_split_EUL_source  src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379
 *   flux_sweep_x/_y    <- flux_generator_x/_y (Transversal = false)
 *                                                          src/flux_calc/flux_generator_x.c:31-165, _y.c:31-165
 *
 * One deliberate deviation: when GRP_2D_flux fails its star/derivative check
 * the reference stores whatever flux the same OpenMP thread produced last (a
 * schedule-dependent value) and then stops after the update.  Here such an
 * interface stores zero flux and zero interface values; the stop step is the
 * same.
 */
#include "orc_grid2d.h"

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
