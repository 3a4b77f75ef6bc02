/*
 * orc_grp2d.c -- the quasi-1D direct Eulerian GRP solver, the 2D GRP flux and
 * the checks around them, restated from the reference.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 *   orc_grp_q1d          <- linear_GRP_solver_Edir_Q1D  src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554
 *   orc_ifvar_check2     <- ifvar_check (dim 2)         src/inter_process/fluid_var_check.c:21-51
 *   orc_star_dire_check2 <- star_dire_check (dim 2)     src/inter_process/fluid_var_check.c:65-131
 *   orc_grp_2d_flux      <- GRP_2D_flux                 src/flux_calc/flux_solver.c:158-237
 *                           (its interface solver: Q1D, the reference's call at :191, or
 *                           G2D, its commented-out alternative at :189; orc_set_flux_solver)
 *
 * Only the single-fluid members are carried (volume/mass fraction terms are
 * identically zero without MULTIFLUID_BASICS and never reach the flux).
 */
#include <math.h>
#include <stddef.h>
#include "grp_oracle.h"

/* Characteristic (upwind) derivative when the t-axis lies outside the fan:
 * the Q1D "trivial" formulas, :149-152 and :266-269 (both sides identical). */
static void side_trivial(double D[4], double U[4], const orc_ifv *S, double c,
                         double lu, double lv)
{
    U[0] = S->rho; U[1] = S->u; U[2] = S->v; U[3] = S->p;
    D[0] = -(S->u - lu) * S->d_rho - (S->v - lv) * S->t_rho - S->rho * (S->d_u + S->t_v);
    D[1] = -(S->u - lu) * S->d_u   - (S->v - lv) * S->t_u   - S->d_p / S->rho;
    D[2] = -(S->u - lu) * S->d_v   - (S->v - lv) * S->t_v   - S->t_p / S->rho;
    D[3] = -(S->u - lu) * S->d_p   - (S->v - lv) * S->t_p   - S->rho * c * c * (S->d_u + S->t_v);
}

void orc_grp_q1d(double wave_speed[2], double D[4], double U[4], double U_star[6],
                 const orc_ifv *L, const orc_ifv *R, double eps, double atc)
{
    const double lu = L->lambda_u, lv = L->lambda_v;
    const double gL = L->gamma, gR = R->gamma;
    const double rho_L = L->rho, rho_R = R->rho;
    const double d_rho_L = L->d_rho, d_rho_R = R->d_rho;
    const double t_rho_L = L->t_rho, t_rho_R = R->t_rho;
    const double u_L = L->u, u_R = R->u;
    const double d_u_L = L->d_u, d_u_R = R->d_u;
    const double t_u_L = L->t_u, t_u_R = R->t_u;
    const double v_L = L->v, v_R = R->v;
    const double d_v_L = L->d_v, d_v_R = R->d_v;
    const double t_v_L = L->t_v, t_v_R = R->t_v;
    const double p_L = L->p, p_R = R->p;
    const double d_p_L = L->d_p, d_p_R = R->d_p;
    const double t_p_L = L->t_p, t_p_R = R->t_p;

    const double zL = (gL - 1.0) / (gL + 1.0);
    const double zR = (gR - 1.0) / (gR + 1.0);

    int CRW[2] = {0, 0};
    double u_star, p_star, rs_L, rs_R, cs_L, cs_R, speed_L, speed_R, C = 0.0;

    const double c_L = sqrt(gL * p_L / rho_L);
    const double c_R = sqrt(gR * p_R / rho_R);
    const double dist = sqrt((rho_L - rho_R) * (rho_L - rho_R) + (u_L - u_R) * (u_L - u_R)
                             + (p_L - p_R) * (p_L - p_R));

    /* star state (:96-134) */
    if (dist < atc && atc < 2 * eps) {
        u_star = 0.5 * (u_R + u_L);
        p_star = 0.5 * (p_R + p_L);
        rs_L = rho_L; cs_L = c_L; speed_L = u_star - cs_L;
        rs_R = rho_R; cs_R = c_R; speed_R = u_star + cs_R;
    } else {
        orc_riemann_exact_rule(&u_star, &p_star, gL, gR, u_L, u_R, p_L, p_R, c_L, c_R, CRW,
                               eps, orc_newton_tol > 0.0 ? orc_newton_tol : eps, 500, 0, orc_newton_rule);
        if (CRW[0]) {
            rs_L = rho_L * pow(p_star / p_L, 1.0 / gL);
            cs_L = c_L * pow(p_star / p_L, 0.5 * (gL - 1.0) / gL);
            speed_L = u_L - c_L;
        } else {
            rs_L = rho_L * (p_star + zL * p_L) / (p_L + zL * p_star);
            cs_L = sqrt(gL * p_star / rs_L);
            speed_L = u_L - c_L * sqrt(0.5 * ((gL + 1.0) * (p_star / p_L) + (gL - 1.0)) / gL);
        }
        if (CRW[1]) {
            rs_R = rho_R * pow(p_star / p_R, 1.0 / gR);
            cs_R = c_R * pow(p_star / p_R, 0.5 * (gR - 1.0) / gR);
            speed_R = u_R + c_R;
        } else {
            rs_R = rho_R * (p_star + zR * p_R) / (p_R + zR * p_star);
            cs_R = sqrt(gR * p_star / rs_R);
            speed_R = u_R + c_R * sqrt(0.5 * ((gR + 1.0) * (p_star / p_R) + (gR - 1.0)) / gR);
        }
    }
    wave_speed[0] = speed_L;
    wave_speed[1] = speed_R;

    U_star[0] = rs_L; U_star[1] = u_star; U_star[2] = rs_R;
    U_star[3] = p_star; U_star[4] = cs_L; U_star[5] = cs_R;

    if (dist < atc) {
        /* acoustic case (:139-254) */
        if (speed_L > lu) { side_trivial(D, U, L, c_L, lu, lv); return; }
        if (speed_R < lu) { side_trivial(D, U, R, c_R, lu, lv); return; }
        if (CRW[0] && ((u_star - cs_L) > lu)) {
            U[1] = zL * (u_L + 2.0 * (c_L + lu) / (gL - 1.0));
            C = U[1] - lu;
            U[3] = pow(C / c_L, 2.0 * gL / (gL - 1.0)) * p_L;
            U[0] = gL * U[3] / C / C;
            U[2] = v_L;
        } else if (CRW[1] && ((u_star + cs_R) < lu)) {
            U[1] = zR * (u_R - 2.0 * (c_R - lu) / (gR - 1.0));
            C = lu - U[1];
            U[3] = pow(C / c_R, 2.0 * gR / (gR - 1.0)) * p_R;
            U[0] = gR * U[3] / C / C;
            U[2] = v_R;
        } else if (u_star > lu) {
            U[0] = rs_L; U[1] = u_star; U[2] = v_L; U[3] = p_star; C = cs_L;
        } else {
            U[0] = rs_R; U[1] = u_star; U[2] = v_R; U[3] = p_star; C = cs_R;
        }
        const double D_p = 0.5 * ((d_u_L * (U[0] * C) + d_p_L) - (d_u_R * (U[0] * C) - d_p_R));
        const double T_p = 0.5 * ((t_u_L * (U[0] * C) + t_p_L) - (t_u_R * (U[0] * C) - t_p_R));
        const double D_u = 0.5 * (d_u_L + d_p_L / (U[0] * C) + d_u_R - d_p_R / (U[0] * C));
        const double T_u = 0.5 * (t_u_L + t_p_L / (U[0] * C) + t_u_R - t_p_R / (U[0] * C));
        double D_v, T_v, D_rho, T_rho;
        if (u_star > lu) {
            D_v = d_v_L; T_v = t_v_L;
            D_rho = d_rho_L - d_p_L / (C * C) + D_p / (C * C);
            T_rho = t_rho_L - t_p_L / (C * C) + T_p / (C * C);
        } else {
            D_v = d_v_R; T_v = t_v_R;
            D_rho = d_rho_R - d_p_R / (C * C) + D_p / (C * C);
            T_rho = t_rho_R - t_p_R / (C * C) + T_p / (C * C);
        }
        D[0] = -(U[1] - lu) * D_rho - (U[2] - lv) * T_rho - U[0] * (D_u + T_v);
        D[1] = -(U[1] - lu) * D_u   - (U[2] - lv) * T_u   - D_p / U[0];
        D[2] = -(U[1] - lu) * D_v   - (U[2] - lv) * T_v   - T_p / U[0];
        D[3] = -(U[1] - lu) * D_p   - (U[2] - lv) * T_p   - U[0] * C * C * (D_u + T_v);
        return;
    }

    /* non-acoustic case (:256-547) */
    if (speed_L > lu) { side_trivial(D, U, L, c_L, lu, lv); return; }
    if (speed_R < lu) { side_trivial(D, U, R, c_R, lu, lv); return; }

    if (CRW[0] && ((u_star - cs_L) > lu)) {            /* t-axis inside the 1-CRW (:290-314) */
        U[1] = zL * (u_L + 2.0 * (c_L + lu) / (gL - 1.0));
        C = U[1] - lu;
        U[3] = pow(C / c_L, 2.0 * gL / (gL - 1.0)) * p_L;
        U[0] = gL * U[3] / C / C;
        U[2] = v_L;
        const double c_frac = C / c_L;
        const double TdS = (d_p_L - d_rho_L * c_L * c_L) / (gL - 1.0) / rho_L;
        const double d_Psi = d_u_L + (gL * d_p_L / c_L - c_L * d_rho_L) / (gL - 1.0) / rho_L;
        D[1] = ((1.0 + zL) * pow(c_frac, 0.5 / zL) + zL * pow(c_frac, (1.0 + zL) / zL));
        D[1] = D[1] / (1.0 + 2.0 * zL) * TdS;
        D[1] = D[1] - c_L * pow(c_frac, 0.5 / zL) * d_Psi;
        D[3] = U[0] * (U[1] - lu) * D[1];
        D[0] = U[0] * (U[1] - lu) * pow(c_frac, (1.0 + zL) / zL) * TdS * (gL - 1.0);
        D[0] = (D[0] + D[3]) / C / C;
        D[2] = -(U[1] - lu) * d_v_L * U[0] / rho_L;
    } else if (CRW[1] && ((u_star + cs_R) < lu)) {     /* t-axis inside the 3-CRW (:315-339) */
        U[1] = zR * (u_R - 2.0 * (c_R - lu) / (gR - 1.0));
        C = lu - U[1];
        U[3] = pow(C / c_R, 2.0 * gR / (gR - 1.0)) * p_R;
        U[0] = gR * U[3] / C / C;
        U[2] = v_R;
        const double c_frac = C / c_R;
        const double TdS = (d_p_R - d_rho_R * c_R * c_R) / (gR - 1.0) / rho_R;
        const double d_Phi = d_u_R - (gR * d_p_R / c_R - c_R * d_rho_R) / (gR - 1.0) / rho_R;
        D[1] = ((1.0 + zR) * pow(c_frac, 0.5 / zR) + zR * pow(c_frac, (1.0 + zR) / zR));
        D[1] = D[1] / (1.0 + 2.0 * zR) * TdS;
        D[1] = D[1] + c_R * pow(c_frac, 0.5 / zR) * d_Phi;
        D[3] = U[0] * (U[1] - lu) * D[1];
        D[0] = U[0] * (U[1] - lu) * pow(c_frac, (1.0 + zR) / zR) * TdS * (gR - 1.0);
        D[0] = (D[0] + D[3]) / C / C;
        D[2] = -(U[1] - lu) * d_v_R * U[0] / rho_R;
    } else {                                           /* non-sonic (:340-523) */
        double a_L, b_L, d_L, a_R, b_R, d_R;
        if (u_star < lu) {
            U[0] = rs_R; U[1] = u_star; U[2] = v_R; U[3] = p_star; C = cs_R;
        } else {
            U[0] = rs_L; U[1] = u_star; U[2] = v_L; U[3] = p_star; C = cs_L;
        }
        if (CRW[0]) {
            a_L = 1.0;
            b_L = 1.0 / rs_L / cs_L;
            const double c_frac = cs_L / c_L;
            const double TdS = (d_p_L - d_rho_L * c_L * c_L) / (gL - 1.0) / rho_L;
            const double d_Psi = d_u_L + (gL * d_p_L / c_L - c_L * d_rho_L) / (gL - 1.0) / rho_L;
            d_L = ((1.0 + zL) * pow(c_frac, 0.5 / zL) + zL * pow(c_frac, (1.0 + zL) / zL));
            d_L = d_L / (1.0 + 2.0 * zL) * TdS;
            d_L = d_L - c_L * pow(c_frac, 0.5 / zL) * d_Psi;
        } else {
            const double SmUs = -sqrt(0.5 * ((gL + 1.0) * p_L + (gL - 1.0) * p_star) / rs_L);
            const double SmUL = -sqrt(0.5 * ((gL + 1.0) * p_star + (gL - 1.0) * p_L) / rho_L);
            const double VAR = sqrt((1 - zL) / (rho_L * (p_star + zL * p_L)));
            const double H1 = 0.5 * VAR * (p_star + (1.0 + 2.0 * zL) * p_L) / (p_star + zL * p_L);
            const double H2 = -0.5 * VAR * ((2.0 + zL) * p_star + zL * p_L) / (p_star + zL * p_L);
            const double H3 = -0.5 * VAR * (p_star - p_L) / rho_L;
            const double L_p = -1.0 / rho_L - SmUL * H2;
            const double L_u = SmUL + rho_L * (c_L * c_L * H2 + H3);
            const double L_rho = -SmUL * H3;
            a_L = 1.0 - rs_L * SmUs * H1;
            b_L = -SmUs / (rs_L * cs_L * cs_L) + H1;
            d_L = L_rho * d_rho_L + L_u * d_u_L + L_p * d_p_L;
        }
        if (CRW[1]) {
            a_R = 1.0;
            b_R = -1.0 / rs_R / cs_R;
            const double c_frac = cs_R / c_R;
            const double TdS = (d_p_R - d_rho_R * c_R * c_R) / (gR - 1.0) / rho_R;
            const double d_Phi = d_u_R - (gR * d_p_R / c_R - c_R * d_rho_R) / (gR - 1.0) / rho_R;
            d_R = ((1.0 + zR) * pow(c_frac, 0.5 / zR) + zR * pow(c_frac, (1.0 + zR) / zR));
            d_R = d_R / (1.0 + 2.0 * zR) * TdS;
            d_R = d_R + c_R * pow(c_frac, 0.5 / zR) * d_Phi;
        } else {
            const double SmUs = sqrt(0.5 * ((gR + 1.0) * p_R + (gR - 1.0) * p_star) / rs_R);
            const double SmUR = sqrt(0.5 * ((gR + 1.0) * p_star + (gR - 1.0) * p_R) / rho_R);
            const double VAR = sqrt((1.0 - zR) / (rho_R * (p_star + zR * p_R)));
            const double H1 = 0.5 * VAR * (p_star + (1 + 2.0 * zR) * p_R) / (p_star + zR * p_R);
            const double H2 = -0.5 * VAR * ((2.0 + zR) * p_star + zR * p_R) / (p_star + zR * p_R);
            const double H3 = -0.5 * VAR * (p_star - p_R) / rho_R;
            const double L_p = -1.0 / rho_R + SmUR * H2;
            const double L_u = SmUR - rho_R * (c_R * c_R * H2 + H3);
            const double L_rho = SmUR * H3;
            a_R = 1.0 + rs_R * SmUs * H1;
            b_R = -(SmUs / (rs_R * cs_R * cs_R) + H1);
            d_R = L_rho * d_rho_R + L_u * d_u_R + L_p * d_p_R;
        }
        const double detA = a_L * b_R - b_L * a_R;
        const double u_t_mat = (b_R * d_L - b_L * d_R) / detA;
        const double p_t_mat = (a_L * d_R - a_R * d_L) / detA;
        D[1] = u_t_mat + (u_star - lu) / U[0] / C / C * p_t_mat;
        D[3] = p_t_mat + (u_star - lu) * U[0] * u_t_mat;

        if (u_star < lu) {
            if (CRW[1]) {
                D[0] = rs_R * (u_star - lu) * pow(cs_R / c_R, (1.0 + zR) / zR) * (d_p_R - d_rho_R * c_R * c_R) / rho_R;
                D[0] = (D[0] + D[3]) / cs_R / cs_R;
                D[2] = -U[1] * d_v_R * U[0] / rho_R;
                D[2] = D[2] + lu * d_v_R;
            } else {
                const double SmUs = sqrt(0.5 * ((gR + 1.0) * p_R + (gR - 1.0) * p_star) / rs_R);
                const double SmUR = sqrt(0.5 * ((gR + 1.0) * p_star + (gR - 1.0) * p_R) / rho_R);
                const double VAR = p_R + zR * p_star;
                const double H1 = rho_R * p_R * (1.0 - zR * zR) / VAR / VAR;
                const double H2 = rho_R * p_star * (zR * zR - 1.0) / VAR / VAR;
                const double H3 = (p_star + zR * p_R) / VAR;
                const double L_rho = SmUR * H3 * d_rho_R;
                const double L_u = -rho_R * (H2 * c_R * c_R + H3) * d_u_R;
                const double L_p = H2 * SmUR * d_p_R;
                D[0] = ((u_star + SmUs) / cs_R / cs_R - u_star * H1) * p_t_mat + rs_R * u_star * SmUs * H1 * u_t_mat;
                D[0] = (D[0] - u_star * (L_p + L_rho + L_u)) / SmUs;
                const double f = SmUR * (H2 * d_p_R + H3 * d_rho_R) - rho_R * (H2 * c_R * c_R + H3) * d_u_R;
                const double rho_x = (f + H1 * (p_t_mat - rs_R * SmUs * u_t_mat) - D[0]) / (SmUR + u_R);
                D[0] = D[0] + lu * rho_x;
                D[2] = -U[1] * SmUR * d_v_R / SmUs;
                D[2] = D[2] + lu * d_v_R;
            }
        } else {
            if (CRW[0]) {
                D[0] = rs_L * (u_star - lu) * pow(cs_L / c_L, (1.0 + zL) / zL) * (d_p_L - d_rho_L * c_L * c_L) / rho_L;
                D[0] = (D[0] + D[3]) / cs_L / cs_L;
                D[2] = -U[1] * d_v_L * U[0] / rho_L;
                D[2] = D[2] + lu * d_v_L;
            } else {
                const double SmUs = -sqrt(0.5 * ((gL + 1.0) * p_L + (gL - 1.0) * p_star) / rs_L);
                const double SmUL = -sqrt(0.5 * ((gL + 1.0) * p_star + (gL - 1.0) * p_L) / rho_L);
                const double VAR = p_L + zL * p_star;
                const double H1 = rho_L * p_L * (1.0 - zL * zL) / VAR / VAR;
                const double H2 = rho_L * p_star * (zL * zL - 1.0) / VAR / VAR;
                const double H3 = (p_star + zL * p_L) / VAR;
                const double L_rho = SmUL * H3 * d_rho_L;
                const double L_u = -rho_L * (H2 * c_L * c_L + H3) * d_u_L;
                const double L_p = H2 * SmUL * d_p_L;
                D[0] = ((u_star + SmUs) / cs_L / cs_L - H1 * u_star) * p_t_mat + rs_L * u_star * SmUs * H1 * u_t_mat;
                D[0] = (D[0] - u_star * (L_p + L_rho + L_u)) / SmUs;
                const double f = SmUL * (H2 * d_p_L + H3 * d_rho_L) - rho_L * (H2 * c_L * c_L + H3) * d_u_L;
                const double rho_x = (f + H1 * (p_t_mat - rs_L * SmUs * u_t_mat) - D[0]) / (SmUL + u_L);
                D[0] = D[0] + lu * rho_x;
                D[2] = -U[1] * SmUL * d_v_L / SmUs;
                D[2] = D[2] + lu * d_v_L;
            }
        }
    }
    /* transversal terms (:524-545) */
    const double T_p = 0.5 * ((t_u_L * (U[0] * C) + t_p_L) - (t_u_R * (U[0] * C) - t_p_R));
    const double T_u = 0.5 * (t_u_L + t_p_L / (U[0] * C) + t_u_R - t_p_R / (U[0] * C));
    if (u_star > lu) {
        const double T_rho = t_rho_L - t_p_L / (C * C) + T_p / (C * C);
        D[0] = D[0] - (U[2] - lv) * T_rho - U[0] * t_v_L;
        D[1] = D[1] - (U[2] - lv) * T_u;
        D[2] = D[2] - (U[2] - lv) * t_v_L - T_p / U[0];
        D[3] = D[3] - (U[2] - lv) * T_p - U[0] * C * C * t_v_L;
    } else {
        const double T_rho = t_rho_R - t_p_R / (C * C) + T_p / (C * C);
        D[0] = D[0] - (U[2] - lv) * T_rho - U[0] * t_v_R;
        D[1] = D[1] - (U[2] - lv) * T_u;
        D[2] = D[2] - (U[2] - lv) * t_v_R - T_p / U[0];
        D[3] = D[3] - (U[2] - lv) * T_p - U[0] * C * C * t_v_R;
    }
}

int orc_ifvar_check2(const orc_ifv *L, const orc_ifv *R, double eps)
{
    if (L->p < eps || R->p < eps || L->rho < eps || R->rho < eps)
        return 1;
    if (!isfinite(L->d_p) || !isfinite(R->d_p) || !isfinite(L->d_u) || !isfinite(R->d_u)
        || !isfinite(L->d_v) || !isfinite(R->d_v) || !isfinite(L->d_rho) || !isfinite(R->d_rho))
        return 2;
    if (!isfinite(L->t_p) || !isfinite(R->t_p) || !isfinite(L->t_u) || !isfinite(R->t_u)
        || !isfinite(L->t_v) || !isfinite(R->t_v) || !isfinite(L->t_rho) || !isfinite(R->t_rho))
        return 2;
    return 0;
}

int orc_star_dire_check2(const double mid[4], const double dire[4], double eps)
{
    if (mid[3] < eps || mid[0] < eps)
        return 1;
    if (!isfinite(mid[1]) || !isfinite(mid[2]) || !isfinite(mid[0]) || !isfinite(mid[3]))
        return 2;
    if (!isfinite(dire[1]) || !isfinite(dire[2]) || !isfinite(dire[0]) || !isfinite(dire[3]))
        return 3;
    return 0;
}

int orc_flux_solver = 0;
void orc_set_flux_solver(int solver) { orc_flux_solver = solver; }

int orc_grp_2d_flux(orc_ifv L, orc_ifv R, double tau, double eps,
                    double n_x, double n_y, double F[4], double I[4])
{
    const double gamma = L.gamma;
    L.lambda_u = 0.0; L.lambda_v = 0.0;

    /* rotation to the interface normal (:167-185) */
    const double u   =  L.u * n_x + L.v * n_y;
    const double u_R =  R.u * n_x + R.v * n_y;
    const double d_u =  L.d_u * n_x + L.d_v * n_y;
    const double d_u_R = R.d_u * n_x + R.d_v * n_y;
    const double t_u =  L.t_u * n_x + L.t_v * n_y;
    const double t_u_R = R.t_u * n_x + R.t_v * n_y;
    L.v   = -L.u * n_y + L.v * n_x;
    R.v   = -R.u * n_y + R.v * n_x;
    L.d_v = -L.d_u * n_y + L.d_v * n_x;
    R.d_v = -R.d_u * n_y + R.d_v * n_x;
    L.t_v = -L.t_u * n_y + L.t_v * n_x;
    R.t_v = -R.t_u * n_y + R.t_v * n_x;
    L.u = u; R.u = u_R; L.d_u = d_u; R.d_u = d_u_R; L.t_u = t_u; R.t_u = t_u_R;

    double wave_speed[2], dire[6], mid[6], star[6];
    if (orc_flux_solver == 1)   /* the reference's alternative call (:189) */
        orc_grp_g2d(wave_speed, dire, mid, star, &L, &R, NULL, NULL, eps, eps);
    else
        orc_grp_q1d(wave_speed, dire, mid, star, &L, &R, eps, eps);

    const int retval = orc_star_dire_check2(mid, dire, eps);
    if (retval)
        return retval;

    const double rho_mid = mid[0] + 0.5 * tau * dire[0];
    const double u_mid = (mid[1] + 0.5 * tau * dire[1]) * n_x - (mid[2] + 0.5 * tau * dire[2]) * n_y;
    const double v_mid = (mid[1] + 0.5 * tau * dire[1]) * n_y + (mid[2] + 0.5 * tau * dire[2]) * n_x;
    const double p_mid = mid[3] + 0.5 * tau * dire[3];

    F[0] = rho_mid * (u_mid * n_x + v_mid * n_y);
    F[1] = F[0] * u_mid + p_mid * n_x;
    F[2] = F[0] * v_mid + p_mid * n_y;
    F[3] = (gamma / (gamma - 1.0)) * p_mid / rho_mid + 0.5 * (u_mid * u_mid + v_mid * v_mid);
    F[3] = F[0] * F[3];

    I[1] = (mid[1] + tau * dire[1]) * n_x - (mid[2] + tau * dire[2]) * n_y;
    I[2] = (mid[1] + tau * dire[1]) * n_y + (mid[2] + tau * dire[2]) * n_x;
    I[0] = mid[0] + tau * dire[0];
    I[3] = mid[3] + tau * dire[3];
    return 0;
}
