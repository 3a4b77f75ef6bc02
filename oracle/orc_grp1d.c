/*
 * orc_grp1d.c -- the 1D Eulerian GRP path, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_grp_edir  <- linear_GRP_solver_Edir   src/riemann_solver/linear_grp_solver_Edir.c:33-337
 *   orc_grp1d     <- GRP_solver_EUL_source    src/finite_volume/grp_solver_EUL_source.c:27-334
 *   bc_limiter_1d <- bound_cond_slope_limiter src/inter_process/bound_cond_slope_limiter.c:30-136
 *                    (fixed h, with minmod_limiter, src/inter_process/slope_limiter.c:31-83)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "grp_oracle.h"

/* the characteristic derivative formulas used on the right side (and on the
 * left in the acoustic case): :69-76 and :131-138 */
static void trivial_upwind(double D[3], double U[3], double rho, double u, double p,
                           double s_rho, double s_u, double s_p, double gamma)
{
    D[0] = -s_rho * u - rho * s_u;
    D[1] = (D[0] * u + s_rho * u * u + 2.0 * rho * u * s_u + s_p) / -rho;
    D[2] = -(gamma - 1.0) * (0.5 * D[0] * u * u + rho * u * D[1]);
    D[2] = D[2] - s_u * (gamma * p + 0.5 * (gamma - 1.0) * rho * u * u);
    D[2] = D[2] - u * (gamma * s_p + (gamma - 1.0) * (0.5 * s_rho * u * u + rho * u * s_u));
    U[0] = rho; U[1] = u; U[2] = p;
}

/* derivative coefficient of a centred rarefaction evaluated at sound-speed ratio cf */
static double crw_coeff(double cf, double zeta)
{
    return 0.5 * (pow(cf, 0.5 / zeta) * (1.0 + zeta) + pow(cf, (1.0 + zeta) / zeta) * zeta) / (0.5 + zeta);
}

void orc_grp_edir(double D[3], double U[3], const orc_ifv *Ls, const orc_ifv *Rs,
                  double eps, double atc)
{
    const double rho_L = Ls->rho, rho_R = Rs->rho;
    const double s_rho_L = Ls->d_rho, s_rho_R = Rs->d_rho;
    const double u_L = Ls->u, u_R = Rs->u;
    const double s_u_L = Ls->d_u, s_u_R = Rs->d_u;
    const double p_L = Ls->p, p_R = Rs->p;
    const double s_p_L = Ls->d_p, s_p_R = Rs->d_p;
    const double gamma = Ls->gamma;
    const double zeta = (gamma - 1.0) / (gamma + 1.0), zts = zeta * zeta;

    int CRW[2] = {0, 0};
    double u_star, p_star, rs_L, rs_R, cs_L, cs_R;

    const double c_L = sqrt(gamma * p_L / rho_L);
    const double c_R = sqrt(gamma * p_R / rho_R);
    const double dist = sqrt((u_L - u_R) * (u_L - u_R) + (p_L - p_R) * (p_L - p_R));

    if (dist < atc && atc < 2 * eps) {
        rs_L = rho_L; rs_R = rho_R; cs_L = c_L; cs_R = c_R;
        u_star = 0.5 * (u_R + u_L);
        p_star = 0.5 * (p_R + p_L);
    } else {
        orc_riemann_exact_ben(&u_star, &p_star, gamma, u_L, u_R, p_L, p_R, c_L, c_R, CRW,
                              eps, eps, 50, 0);
        rs_L = (p_star > p_L) ? rho_L * (p_star + zeta * p_L) / (p_L + zeta * p_star)
                              : rho_L * pow(p_star / p_L, 1.0 / gamma);
        rs_R = (p_star > p_R) ? rho_R * (p_star + zeta * p_R) / (p_R + zeta * p_star)
                              : rho_R * pow(p_star / p_R, 1.0 / gamma);
        cs_L = sqrt(gamma * p_star / rs_L);
        cs_R = sqrt(gamma * p_star / rs_R);
    }

    if (dist < atc) {                                                  /* :79-139 */
        if (u_L - c_L > 0.0) { trivial_upwind(D, U, rho_L, u_L, p_L, s_rho_L, s_u_L, s_p_L, gamma); return; }
        if (u_R + c_R < 0.0) { trivial_upwind(D, U, rho_R, u_R, p_R, s_rho_R, s_u_R, s_p_R, gamma); return; }
        const double PI = (u_star + cs_R) * rs_L * cs_L * cs_L / (u_star - cs_L) / rs_R / cs_R / cs_R;
        D[1] = (s_p_L / rho_L + c_L * s_u_L) * PI / (1.0 - PI) + (s_p_R / rho_R - c_R * s_u_R) / (PI - 1.0);
        D[2] = ((u_star + cs_R) / rs_R / cs_R / cs_R) - ((u_star - cs_L) / rs_L / cs_L / cs_L);
        D[2] = (s_p_R / rho_R - c_R * s_u_R - s_p_L / rho_L - c_L * s_u_L) / D[2];
        if (u_star > 0.0) {
            U[0] = rs_L; U[1] = u_star; U[2] = p_star;
            D[2] = D[2] * (1.0 - (u_star * u_star / cs_L / cs_L)) + rs_L * u_star * D[1];
            D[0] = (u_star * (s_p_L - s_rho_L * cs_L * cs_L) + D[2]) / cs_L / cs_L;
        } else {
            U[0] = rs_R; U[1] = u_star; U[2] = p_star;
            D[2] = D[2] * (1.0 - (u_star * u_star / cs_R / cs_R)) + rs_R * u_star * D[1];
            D[0] = (u_star * (s_p_R - s_rho_R * cs_R * cs_R) + D[2]) / cs_R / cs_R;
        }
        return;
    }

    const double speed_L = CRW[0] ? u_L - c_L : (rs_L * u_star - rho_L * u_L) / (rs_L - rho_L);
    const double speed_R = CRW[1] ? u_R + c_R : (rs_R * u_star - rho_R * u_R) / (rs_R - rho_R);

    if (speed_L > 0.0) {                                               /* :158-171 */
        D[0] = -s_rho_L * u_L - rho_L * s_u_L;
        D[1] = (D[0] * u_L + s_rho_L * u_L * u_L + 2.0 * rho_L * u_L * s_u_L + s_p_L) / -rho_L;
        D[2] = (s_u_L * p_L + u_L * s_p_L) * gamma / (1.0 - gamma) - 0.5 * s_rho_L * u_L * u_L * u_L
               - 1.5 * rho_L * u_L * u_L * s_u_L;
        D[2] = D[2] - 0.5 * D[0] * u_L * u_L - rho_L * u_L * D[1];
        D[2] = D[2] * (gamma - 1.0);
        U[0] = rho_L; U[1] = u_L; U[2] = p_L;
        return;
    }
    if (speed_R < 0.0) { trivial_upwind(D, U, rho_R, u_R, p_R, s_rho_R, s_u_R, s_p_R, gamma); return; }

    if (CRW[0] && ((u_star - cs_L) > 0.0)) {                             /* 1-CRW sonic :188-205 */
        U[1] = zeta * (u_L + 2.0 * c_L / (gamma - 1.0));
        U[2] = U[1] * U[1] * rho_L / gamma / pow(p_L, 1.0 / gamma);
        U[2] = pow(U[2], gamma / (gamma - 1.0));
        U[0] = gamma * U[2] / U[1] / U[1];
        D[1] = crw_coeff(U[1] / c_L, zeta);
        D[1] = D[1] * (s_p_L - s_rho_L * c_L * c_L) / (gamma - 1.0) / rho_L;
        D[1] = D[1] - c_L * pow(U[1] / c_L, 0.5 / zeta) * (s_u_L + (gamma * s_p_L / c_L - c_L * s_rho_L) / (gamma - 1.0) / rho_L);
        D[2] = U[0] * U[1] * D[1];
        D[0] = U[0] * U[1] * pow(U[1] / c_L, (1.0 + zeta) / zeta) * (s_p_L - s_rho_L * c_L * c_L) / rho_L;
        D[0] = (D[0] + D[2]) / U[1] / U[1];
        return;
    }
    if (CRW[1] && ((u_star + cs_R) < 0.0)) {                             /* 3-CRW sonic :206-223 */
        U[1] = zeta * (u_R - 2.0 * c_R / (gamma - 1.0));
        U[2] = U[1] * U[1] * rho_R / gamma / pow(p_R, 1.0 / gamma);
        U[2] = pow(U[2], gamma / (gamma - 1.0));
        U[0] = gamma * U[2] / U[1] / U[1];
        D[1] = crw_coeff(-U[1] / c_R, zeta);
        D[1] = D[1] * (s_p_R - s_rho_R * c_R * c_R) / (gamma - 1.0) / rho_R;
        D[1] = D[1] + c_R * pow(-U[1] / c_R, 0.5 / zeta) * (s_u_R - (gamma * s_p_R / c_R - c_R * s_rho_R) / (gamma - 1.0) / rho_R);
        D[2] = U[0] * U[1] * D[1];
        D[0] = U[0] * U[1] * pow(-U[1] / c_R, (1.0 + zeta) / zeta) * (s_p_R - s_rho_R * c_R * c_R) / rho_R;
        D[0] = (D[0] + D[2]) / U[1] / U[1];
        return;
    }

    double a_L, b_L, d_L, a_R, b_R, d_R;                               /* non-sonic :225-327 */
    if (CRW[0]) {
        a_L = 1.0;
        b_L = 1.0 / rs_L / cs_L;
        d_L = crw_coeff(cs_L / c_L, zeta);
        d_L = d_L * (s_p_L - s_rho_L * c_L * c_L) / (gamma - 1.0) / rho_L;
        d_L = d_L - c_L * pow(cs_L / c_L, 0.5 / zeta) * (s_u_L + (gamma * s_p_L / c_L - c_L * s_rho_L) / (gamma - 1.0) / rho_L);
    } else {
        const double H1 = 0.5 * sqrt((1.0 - zeta) / (rho_L * (p_star + zeta * p_L))) * (p_star + (1.0 + 2.0 * zeta) * p_L) / (p_star + zeta * p_L);
        const double H2 = -0.5 * sqrt((1.0 - zeta) / (rho_L * (p_star + zeta * p_L))) * ((2.0 + zeta) * p_star + zeta * p_L) / (p_star + zeta * p_L);
        const double H3 = -0.5 * sqrt((1.0 - zeta) / (rho_L * (p_star + zeta * p_L))) * (p_star - p_L) / rho_L;
        const double shk = (rs_L * u_star - rho_L * u_L) / (rs_L - rho_L);
        a_L = 1.0 - rs_L * (shk - u_star) * H1;
        b_L = (u_star - shk) / rs_L / cs_L / cs_L + H1;
        const double L_rho = (u_L - shk) * H3;
        const double L_u = shk - u_L + rho_L * c_L * c_L * H2 + rho_L * H3;
        const double L_p = (u_L - shk) * H2 - 1.0 / rho_L;
        d_L = L_rho * s_rho_L + L_u * s_u_L + L_p * s_p_L;
    }
    if (CRW[1]) {
        a_R = 1.0;
        b_R = -1.0 / rs_R / cs_R;
        d_R = crw_coeff(cs_R / c_R, zeta);
        d_R = d_R * (s_p_R - s_rho_R * c_R * c_R) / (gamma - 1.0) / rho_R;
        d_R = d_R + c_R * pow(cs_R / c_R, 0.5 / zeta) * (s_u_R - (gamma * s_p_R / c_R - c_R * s_rho_R) / (gamma - 1.0) / rho_R);
    } else {
        const double H1 = 0.5 * sqrt((1.0 - zeta) / (rho_R * (p_star + zeta * p_R))) * (p_star + (1.0 + 2.0 * zeta) * p_R) / (p_star + zeta * p_R);
        const double H2 = -0.5 * sqrt((1.0 - zeta) / (rho_R * (p_star + zeta * p_R))) * ((2.0 + zeta) * p_star + zeta * p_R) / (p_star + zeta * p_R);
        const double H3 = -0.5 * sqrt((1.0 - zeta) / (rho_R * (p_star + zeta * p_R))) * (p_star - p_R) / rho_R;
        const double shk = (rs_R * u_star - rho_R * u_R) / (rs_R - rho_R);
        a_R = 1.0 + rs_R * (shk - u_star) * H1;
        b_R = (u_star - shk) / rs_R / cs_R / cs_R - H1;
        const double L_rho = (shk - u_R) * H3;
        const double L_u = shk - u_R - rho_R * c_R * c_R * H2 - rho_R * H3;
        const double L_p = (shk - u_R) * H2 - 1.0 / rho_R;
        d_R = L_rho * s_rho_R + L_u * s_u_R + L_p * s_p_R;
    }
    const double p_t_mat = (d_L * a_R / a_L - d_R) / (b_L * a_R / a_L - b_R);
    const double u_t_mat = (d_L - b_L * p_t_mat) / a_L;

    if (u_star < 0.0) {
        U[0] = rs_R; U[1] = u_star; U[2] = p_star;
        D[1] = u_t_mat + u_star * p_t_mat / rs_R / cs_R / cs_R;
        D[2] = p_t_mat + rs_R * u_star * u_t_mat;
        if (CRW[1]) {
            D[0] = rs_R * u_star * pow(cs_R / c_R, (1.0 + zeta) / zeta) * (s_p_R - s_rho_R * c_R * c_R) / rho_R;
            D[0] = (D[0] + D[2]) / cs_R / cs_R;
        } else {
            const double shk = (rs_R * u_star - rho_R * u_R) / (rs_R - rho_R);
            const double H1 = rho_R * p_R * (1.0 - zts) / (p_R + zeta * p_star) / (p_R + zeta * p_star);
            const double H2 = rho_R * p_star * (zts - 1.0) / (p_R + zeta * p_star) / (p_R + zeta * p_star);
            const double H3 = (p_star + zeta * p_R) / (p_R + zeta * p_star);
            const double g_rho = u_star - shk;
            const double g_u = u_star * rs_R * (shk - u_star) * H1;
            const double g_p = shk / cs_R / cs_R - u_star * H1;
            const double f = (shk - u_R) * (H2 * s_p_R + H3 * s_rho_R) - rho_R * (H2 * c_R * c_R + H3) * s_u_R;
            D[0] = (f * u_star - g_p * p_t_mat - g_u * u_t_mat) / g_rho;
        }
    } else {
        U[0] = rs_L; U[1] = u_star; U[2] = p_star;
        D[1] = u_t_mat + u_star * p_t_mat / rs_L / cs_L / cs_L;
        D[2] = p_t_mat + rs_L * u_star * u_t_mat;
        if (CRW[0]) {
            D[0] = rs_L * u_star * pow(cs_L / c_L, (1.0 + zeta) / zeta) * (s_p_L - s_rho_L * c_L * c_L) / rho_L;
            D[0] = (D[0] + D[2]) / cs_L / cs_L;
        } else {
            const double shk = (rs_L * u_star - rho_L * u_L) / (rs_L - rho_L);
            const double H1 = rho_L * p_L * (1.0 - zts) / (p_L + zeta * p_star) / (p_L + zeta * p_star);
            const double H2 = rho_L * p_star * (zts - 1.0) / (p_L + zeta * p_star) / (p_L + zeta * p_star);
            const double H3 = (p_star + zeta * p_L) / (p_L + zeta * p_star);
            const double g_rho = u_star - shk;
            const double g_u = u_star * rs_L * (shk - u_star) * H1;
            const double g_p = shk / cs_L / cs_L - u_star * H1;
            const double f = (shk - u_L) * (H2 * s_p_L + H3 * s_rho_L) - rho_L * (H2 * c_L * c_L + H3) * s_u_L;
            D[0] = (f * u_star - g_p * p_t_mat - g_u * u_t_mat) / g_rho;
        }
    }
}

typedef struct ghost1 { double RHO, P, U, SRHO, SP, SU; } ghost1;

static void limiter_1d(int m, const double *U, double *s, double UL, double UR, double h,
                       int have_prev, double alpha)
{
    for (int j = 0; j < m; ++j) {
        const double s_L = j ? (U[j] - U[j - 1]) / h : (U[j] - UL) / h;
        const double s_R = (j < m - 1) ? (U[j + 1] - U[j]) / h : (UR - U[j]) / h;
        s[j] = have_prev ? orc_minmod3(alpha * s_L, alpha * s_R, s[j]) : orc_minmod2(s_L, s_R);
    }
}

static int bc_limiter_1d(int m, int bc, const double *rho0, const double *u0, const double *p0,
                         const double *rho, const double *u, const double *p,
                         double *s_rho, double *s_u, double *s_p,
                         ghost1 *L, ghost1 *R, int found, double h, double alpha)
{
    switch (bc) {
    case -1:
        if (found) break;
        L->U = u0[0]; R->U = u0[m - 1];
        L->P = p0[0]; R->P = p0[m - 1];
        L->RHO = rho0[0]; R->RHO = rho0[m - 1];
        break;
    case -2:
        L->U = -u[0]; R->U = -u[m - 1];
        L->P = p[0]; R->P = p[m - 1]; L->RHO = rho[0]; R->RHO = rho[m - 1];
        break;
    case -4:
        L->U = u[0]; R->U = u[m - 1];
        L->P = p[0]; R->P = p[m - 1]; L->RHO = rho[0]; R->RHO = rho[m - 1];
        break;
    case -7:
        L->U = u[m - 1]; R->U = u[0];
        L->P = p[m - 1]; R->P = p[0]; L->RHO = rho[m - 1]; R->RHO = rho[0];
        break;
    case -24:
        L->U = -u[0]; R->U = u[m - 1];
        L->P = p[0]; R->P = p[m - 1]; L->RHO = rho[0]; R->RHO = rho[m - 1];
        break;
    default:
        printf("No suitable boundary coditions in x direction!\n");
        return 0;
    }
    limiter_1d(m, u, s_u, L->U, R->U, h, found, alpha);
    limiter_1d(m, p, s_p, L->P, R->P, h, found, alpha);
    limiter_1d(m, rho, s_rho, L->RHO, R->RHO, h, found, alpha);
    switch (bc) {
    case -2:  L->SU = s_u[0]; R->SU = s_u[m - 1]; break;
    case -7:  L->SU = s_u[m - 1]; R->SU = s_u[0];
              L->SP = s_p[m - 1]; R->SP = s_p[0];
              L->SRHO = s_rho[m - 1]; R->SRHO = s_rho[0]; break;
    case -24: L->SU = s_u[0]; break;
    }
    return 1;
}

int orc_grp1d(int m, double *const *RHO, double *const *U, double *const *P,
              double *const *E, int *N_plot, double *time_plot, double *cpu_time,
              double *config)
{
    const double t_all = config[1], eps = config[4], gamma = config[6], CFL = config[7];
    const int N = (int)config[5];
    const double h = config[10], alpha = config[41];
    const int bc = (int)config[17];
    double tau = config[16];

    int k = 0, nt = 0, found = 0, stop_t = 0, status = 0;
    double time_c = 0.0, cpu_sum = 0.0;
    ghost1 gL = {0, 0, 0, 0, 0, 0}, gR = gL;
    double *s_rho = calloc((size_t)m, sizeof(double)), *s_u = calloc((size_t)m, sizeof(double));
    double *s_p = calloc((size_t)m, sizeof(double));
    double *buf = malloc((size_t)(m + 1) * 9 * sizeof(double));
    if (!s_rho || !s_u || !s_p || !buf) { status = 1; goto done; }
    double *RHO_n = buf, *U_n = buf + (m + 1), *P_n = buf + 2 * (m + 1);
    double *RHO_t = buf + 3 * (m + 1), *U_t = buf + 4 * (m + 1), *P_t = buf + 5 * (m + 1);
    double *F_rho = buf + 6 * (m + 1), *F_u = buf + 7 * (m + 1), *F_e = buf + 8 * (m + 1);

    for (k = 1; k <= N; ++k) {
        const clock_t tic = clock();
        if (time_c >= time_plot[nt] && nt < (*N_plot - 1)) {             /* :115-125 */
            memcpy(RHO[nt + 1], RHO[nt], (size_t)m * sizeof(double));
            memcpy(U[nt + 1], U[nt], (size_t)m * sizeof(double));
            memcpy(E[nt + 1], E[nt], (size_t)m * sizeof(double));
            memcpy(P[nt + 1], P[nt], (size_t)m * sizeof(double));
            nt++;
        }
        double *rho = RHO[nt], *u = U[nt], *p = P[nt], *e = E[nt];
        double h_S_max = INFINITY;

        if (!bc_limiter_1d(m, bc, RHO[0], U[0], P[0], rho, u, p, s_rho, s_u, s_p, &gL, &gR,
                           found, h, alpha)) { status = 1; goto done; }
        found = 1;

        for (int j = 0; j <= m; ++j) {                                    /* :133-214 */
            orc_ifv L, R;
            memset(&L, 0, sizeof L); memset(&R, 0, sizeof R);
            L.gamma = gamma; R.gamma = gamma;
            if (j) {
                L.rho = rho[j - 1] + 0.5 * h * s_rho[j - 1];
                L.u = u[j - 1] + 0.5 * h * s_u[j - 1];
                L.p = p[j - 1] + 0.5 * h * s_p[j - 1];
            } else {
                L.rho = gL.RHO + 0.5 * h * gL.SRHO;
                L.u = gL.U + 0.5 * h * gL.SU;
                L.p = gL.P + 0.5 * h * gL.SP;
            }
            if (j < m) {
                R.rho = rho[j] - 0.5 * h * s_rho[j];
                R.u = u[j] - 0.5 * h * s_u[j];
                R.p = p[j] - 0.5 * h * s_p[j];
            } else {
                R.rho = gR.RHO + 0.5 * h * gR.SRHO;
                R.u = gR.U + 0.5 * h * gR.SU;
                R.p = gR.P + 0.5 * h * gR.SP;
            }
            const double c_L = sqrt(gamma * L.p / L.rho);
            const double c_R = sqrt(gamma * R.p / R.rho);
            h_S_max = fmin(h_S_max, h / (fabs(L.u) + fabs(c_L)));
            h_S_max = fmin(h_S_max, h / (fabs(R.u) + fabs(c_R)));
            if (j) { L.d_u = s_u[j - 1]; L.d_p = s_p[j - 1]; L.d_rho = s_rho[j - 1]; }
            else   { L.d_rho = gL.SRHO; L.d_u = gL.SU; L.d_p = gL.SP; }
            if (j < m) { R.d_u = s_u[j]; R.d_p = s_p[j]; R.d_rho = s_rho[j]; }
            else       { R.d_rho = gR.SRHO; R.d_u = gR.SU; R.d_p = gR.SP; }

            if (L.p < eps || R.p < eps || L.rho < eps || R.rho < eps
                || !isfinite(L.d_p) || !isfinite(R.d_p) || !isfinite(L.d_u) || !isfinite(R.d_u)
                || !isfinite(L.d_rho) || !isfinite(R.d_rho)) {
                printf("reconstruction error on [%d, %d] (t_n, x).\n", k, j);
                status = 1;
                goto done;
            }
            double dire[3], mid[3];
            orc_grp_edir(dire, mid, &L, &R, eps, eps);
            if (mid[2] < eps || mid[0] < eps || !isfinite(mid[1]) || !isfinite(mid[2]) || !isfinite(mid[0])
                || !isfinite(dire[1]) || !isfinite(dire[2]) || !isfinite(dire[0])) {
                printf("star/derivative error on [%d, %d] (t_n, x).\n", k, j);
                stop_t = 1;
            }
            RHO_n[j] = mid[0]; U_n[j] = mid[1]; P_n[j] = mid[2];
            RHO_t[j] = dire[0]; U_t[j] = dire[1]; P_t[j] = dire[2];
        }

        if (isfinite(t_all) || !isfinite(config[16]) || config[16] <= 0.0) {   /* :218-233 */
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
        const double nu = tau / h;
        for (int j = 0; j <= m; ++j) {                                    /* :236-250 */
            RHO_n[j] += 0.5 * tau * RHO_t[j];
            U_n[j] += 0.5 * tau * U_t[j];
            P_n[j] += 0.5 * tau * P_t[j];
            F_rho[j] = RHO_n[j] * U_n[j];
            F_u[j] = F_rho[j] * U_n[j] + P_n[j];
            F_e[j] = (gamma / (gamma - 1.0)) * P_n[j] + 0.5 * F_rho[j] * U_n[j];
            F_e[j] = F_e[j] * U_n[j];
            RHO_n[j] += 0.5 * tau * RHO_t[j];
            U_n[j] += 0.5 * tau * U_t[j];
            P_n[j] += 0.5 * tau * P_t[j];
        }
        for (int j = 0; j < m; ++j) {                                     /* :253-277 */
            const double Mom = rho[j] * u[j] - nu * (F_u[j + 1] - F_u[j]);
            const double Ene = rho[j] * e[j] - nu * (F_e[j + 1] - F_e[j]);
            rho[j] = rho[j] - nu * (F_rho[j + 1] - F_rho[j]);
            u[j] = Mom / rho[j];
            e[j] = Ene / rho[j];
            p[j] = (Ene - 0.5 * Mom * u[j]) * (gamma - 1.0);
            if (p[j] < eps || rho[j] < eps) {
                printf("<0.0 error on [%d, %d] (t_n, x) - Update\n", k, j);
                stop_t = 1;
            }
            s_u[j] = (U_n[j + 1] - U_n[j]) / h;
            s_p[j] = (P_n[j + 1] - P_n[j]) / h;
            s_rho[j] = (RHO_n[j + 1] - RHO_n[j]) / h;
        }
        time_c += tau;
        if (stop_t || time_c > (t_all - eps) || !isfinite(time_c))
            break;
        cpu_sum += ((double)clock() - (double)tic) / (double)CLOCKS_PER_SEC;
        if (cpu_time) cpu_time[nt] = cpu_sum;
    }
    if (stop_t) status = 2;

done:
    config[5] = (double)k;
    *N_plot = nt + 1;
    if (isfinite(time_c))      time_plot[nt] = time_c;
    else if (isfinite(t_all))  time_plot[nt] = t_all;
    else if (isfinite(tau))    time_plot[nt] = k * tau;
    free(s_rho); free(s_u); free(s_p); free(buf);
    return status;
}
