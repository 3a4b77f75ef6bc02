/*
 * orc_unsplit2d.c -- the genuinely 2D (unsplit) Eulerian GRP time loop, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_unsplit2d      <- GRP_solver_2D_EUL_source        src/finite_volume/grp_solver_2D_EUL_source.c:75-300
 *   flux_t_x / flux_t_y <- flux_generator_x/_y (Transversal = true)
 *                                                          src/flux_calc/flux_generator_x.c:31-165, _y.c:31-165
 *
 * Per step: both slope limiters (with the ghost-slope rules across the other
 * direction), x- and y-interface fluxes over tau with the transversal
 * derivatives of the cells beside each interface, one combined conservative
 * update, and the slope refresh from the interface values at t + tau.
 *
 * As in orc_split2d.c, an interface whose GRP star/derivative check fails
 * stores zero flux and zero interface values (the reference stores a
 * schedule-dependent leftover); the step that stops is the same.
 */
#include "orc_grid2d.h"

/* interface state with explicit transversal derivatives t[4] = (rho, u, v, p) */
static void side_state_t(orc_ifv *f, double gamma, double rho, double u, double v, double p,
                         double sr, double su, double sv, double sp, double half_h_sign, const double t[4])
{
    side_state(f, gamma, rho, u, v, p, sr, su, sv, sp, half_h_sign, 0.0);
    f->t_rho = t[0]; f->t_u = t[1]; f->t_v = t[2]; f->t_p = t[3];
}

/* flux_generator_x, Transversal = true: t_* of the cell beside the interface,
 * bfv_L/R.T* at the ends (flux_generator_x.c:96-121) */
static int flux_t_x(grid2d *g, double tau)
{
    const int m = g->m, n = g->n;
    const double hh = 0.5 * g->h_x;
    long long last = -1;
#pragma omp parallel for collapse(2) schedule(static) reduction(max:last)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= m; ++j) {
            int code = 0;
            orc_ifv L, R;
            double tl[4], tr[4];
            if (j) {
                const size_t c = AT(j - 1, i);
                for (int q = 0; q < 4; ++q) tl[q] = g->t[q][c];
                side_state_t(&L, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                             g->s[0][c], g->s[1][c], g->s[2][c], g->s[3][c], hh, tl);
            } else {
                const ghost *b = &g->gL[i];
                tl[0] = b->TRHO; tl[1] = b->TU; tl[2] = b->TV; tl[3] = b->TP;
                side_state_t(&L, g->gamma, b->RHO, b->U, b->V, b->P, b->SRHO, b->SU, b->SV, b->SP, hh, tl);
            }
            if (j < m) {
                const size_t c = AT(j, i);
                for (int q = 0; q < 4; ++q) tr[q] = g->t[q][c];
                side_state_t(&R, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                             g->s[0][c], g->s[1][c], g->s[2][c], g->s[3][c], -hh, tr);
            } else {
                const ghost *b = &g->gR[i];
                tr[0] = b->TRHO; tr[1] = b->TU; tr[2] = b->TV; tr[3] = b->TP;
                side_state_t(&R, g->gamma, b->RHO, b->U, b->V, b->P, b->SRHO, b->SU, b->SV, b->SP, -hh, tr);
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

/* flux_generator_y, Transversal = true: the tangential direction of a
 * y-interface is -x, so t_* = -s_* of the cell beside it, -bfv_D/U.S* at the
 * ends (flux_generator_y.c:96-121) */
static int flux_t_y(grid2d *g, double tau)
{
    const int m = g->m, n = g->n;
    const double hh = 0.5 * g->h_y;
    long long last = -1;
#pragma omp parallel for collapse(2) schedule(static) reduction(max:last)
    for (int j = 0; j < m; ++j)
        for (int i = 0; i <= n; ++i) {
            int code = 0;
            orc_ifv D, U;
            double td[4], tu[4];
            if (i) {
                const size_t c = AT(j, i - 1);
                for (int q = 0; q < 4; ++q) td[q] = -g->s[q][c];
                side_state_t(&D, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                             g->t[0][c], g->t[1][c], g->t[2][c], g->t[3][c], hh, td);
            } else {
                const ghost *b = &g->gD[j];
                td[0] = -b->SRHO; td[1] = -b->SU; td[2] = -b->SV; td[3] = -b->SP;
                side_state_t(&D, g->gamma, b->RHO, b->U, b->V, b->P, b->TRHO, b->TU, b->TV, b->TP, hh, td);
            }
            if (i < n) {
                const size_t c = AT(j, i);
                for (int q = 0; q < 4; ++q) tu[q] = -g->s[q][c];
                side_state_t(&U, g->gamma, g->rho[c], g->u[c], g->v[c], g->p[c],
                             g->t[0][c], g->t[1][c], g->t[2][c], g->t[3][c], -hh, tu);
            } else {
                const ghost *b = &g->gU[j];
                tu[0] = -b->SRHO; tu[1] = -b->SU; tu[2] = -b->SV; tu[3] = -b->SP;
                side_state_t(&U, g->gamma, b->RHO, b->U, b->V, b->P, b->TRHO, b->TU, b->TV, b->TP, -hh, tu);
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

/* combined update and refresh of both slope sets (grp_solver_2D_EUL_source.c:231-262);
 * returns 1 if positivity failed */
static int update_xy(grid2d *g, double nu, double mu)
{
    const int m = g->m, n = g->n;
    const double gm1 = g->gamma - 1.0;
    int bad = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(|:bad)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            const size_t c = AT(j, i);
            const size_t x0 = (size_t)j * n + i, x1 = (size_t)(j + 1) * n + i;
            const size_t y0 = (size_t)j * (n + 1) + i, y1 = y0 + 1;
            const double mom_x = g->rho[c] * g->u[c] - nu * (g->Fx[1][x1] - g->Fx[1][x0]) - mu * (g->Gy[1][y1] - g->Gy[1][y0]);
            const double mom_y = g->rho[c] * g->v[c] - nu * (g->Fx[2][x1] - g->Fx[2][x0]) - mu * (g->Gy[2][y1] - g->Gy[2][y0]);
            const double ene   = g->rho[c] * g->e[c] - nu * (g->Fx[3][x1] - g->Fx[3][x0]) - mu * (g->Gy[3][y1] - g->Gy[3][y0]);
            g->rho[c] += -nu * (g->Fx[0][x1] - g->Fx[0][x0]) - mu * (g->Gy[0][y1] - g->Gy[0][y0]);
            g->u[c] = mom_x / g->rho[c];
            g->v[c] = mom_y / g->rho[c];
            g->e[c] = ene / g->rho[c];
            g->p[c] = (ene - 0.5 * mom_x * g->u[c] - 0.5 * mom_y * g->v[c]) * gm1;
            if (g->p[c] < g->eps || g->rho[c] < g->eps)
                bad = 1;
            for (int q = 0; q < 4; ++q) {
                g->s[q][c] = (g->Ix[q][x1] - g->Ix[q][x0]) / g->h_x;
                g->t[q][c] = (g->Iy[q][y1] - g->Iy[q][y0]) / g->h_y;
            }
        }
    return bad;
}

int orc_unsplit2d(int m, int n, double *const *rho, double *const *u, double *const *v,
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

    int k = 0, nt = 0, nt_plot = 0, found_x = 0, found_y = 0, stop_t = 0;
    double time_c = 0.0, nu = 0.0, mu = 0.0, cpu_sum = 0.0;
    if (status) { printf("NOT enough memory!\n"); goto done; }

    for (k = 1; k <= N; ++k) {
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

        /* CFL bound from CV[0] every step (:183-207): frame 0 is the live state
         * until the first plot frame has been copied off */
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
        nu = tau / h_x;
        mu = tau / h_y;

        if (!bc_limiter_x(g, found_x)) { status = 1; goto done; }          /* :211-216 */
        found_x = 1;
        if (!bc_limiter_y(g, found_y)) { status = 1; goto done; }
        found_y = 1;

        int ferr = flux_t_x(g, tau);                                         /* :218-227 */
        if (ferr == 1) { status = 1; goto done; }
        if (ferr == 2) stop_t = 1;
        ferr = flux_t_y(g, tau);
        if (ferr == 1) { status = 1; goto done; }
        if (ferr == 2) stop_t = 1;

        if (update_xy(g, nu, mu)) { printf("<0.0 error on step %d - Update\n", k); stop_t = 1; }

        time_c += tau;
        if (stop_t || time_c > (t_all - eps) || !isfinite(time_c))
            break;
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
