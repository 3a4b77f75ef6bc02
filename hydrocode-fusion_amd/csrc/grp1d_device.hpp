// grp1d_device.hpp -- device form of the 1D direct Eulerian GRP solver.
//
//   grp_edir <- linear_GRP_solver_Edir  src/riemann_solver/linear_grp_solver_Edir.c:33-337
//
// with the exact star state from Riemann_solver_exact_Ben
// (src/riemann_solver/riemann_solver_exact_Ben.c:230-408; the single-gamma
// form of riemann_exact, at most 50 Newton steps, tol = eps).  Expression order
// follows the reference so that, with -ffp-contract=off, results agree with
// its CPU build to the last few ulps.  Powers of one base share its logarithm
// (PowBase).
#pragma once
#include "grp_device.hpp"

namespace hgrp {

// the characteristic derivative formulas of a uniform side (:69-76, :131-138)
__device__ __forceinline__ void edir_upwind(double D[3], double U[3], double rho, double u, double p, double s_rho,
                                            double s_u, double s_p, const Gas &G)
{
    const double gamma = G.g;
    D[0] = -s_rho * u - rho * s_u;
    D[1] = (D[0] * u + s_rho * u * u + 2.0 * rho * u * s_u + s_p) / -rho;
    D[2] = -(gamma - 1.0) * (0.5 * D[0] * u * u + rho * u * D[1]);
    D[2] = D[2] - s_u * (gamma * p + 0.5 * (gamma - 1.0) * rho * u * u);
    D[2] = D[2] - u * (gamma * s_p + (gamma - 1.0) * (0.5 * s_rho * u * u + rho * u * s_u));
    U[0] = rho; U[1] = u; U[2] = p;
}

// 0.5 (cf^(0.5/z) (1+z) + cf^((1+z)/z) z) / (0.5+z) from the two powers of cf
__device__ __forceinline__ double crw_coeff(double pw_a, double pw_b, double zeta)
{
    return 0.5 * (pw_a * (1.0 + zeta) + pw_b * zeta) / (0.5 + zeta);
}

// Side of the interface: value and x-derivative of rho, u, p.
struct Side1 {
    double rho, u, p, s_rho, s_u, s_p;
};

// linear_GRP_solver_Edir with atc = eps: D = time derivatives, U = values of
// (rho, u, p) on the t-axis.
__device__ inline void grp_edir(double D[3], double U[3], const Side1 &L, const Side1 &R, const Gas &G, double eps,
                                double atc)
{
    const double rho_L = L.rho, rho_R = R.rho, u_L = L.u, u_R = R.u, p_L = L.p, p_R = R.p;
    const double s_rho_L = L.s_rho, s_rho_R = R.s_rho, s_u_L = L.s_u, s_u_R = R.s_u, s_p_L = L.s_p, s_p_R = R.s_p;
    const double gamma = G.g, zeta = G.z, zts = G.zz();

    int crw0 = 0, crw1 = 0;
    double u_star, p_star, rs_L, rs_R, cs_L, cs_R;
    const double c_L = sqrt(gamma * p_L / rho_L);
    const double c_R = sqrt(gamma * p_R / rho_R);
    const double dist = sqrt((u_L - u_R) * (u_L - u_R) + (p_L - p_R) * (p_L - p_R));

    if (dist < atc && atc < 2 * eps) {
        rs_L = rho_L; rs_R = rho_R; cs_L = c_L; cs_R = c_R;
        u_star = 0.5 * (u_R + u_L);
        p_star = 0.5 * (p_R + p_L);
    } else {
        const StarResult st = riemann_exact(G, G, u_L, u_R, p_L, p_R, c_L, c_R, eps, eps, 50);
        u_star = st.u_star; p_star = st.p_star; crw0 = st.crw0; crw1 = st.crw1;
        rs_L = (p_star > p_L) ? rho_L * (p_star + zeta * p_L) / (p_L + zeta * p_star)
                              : rho_L * gpow(p_star / p_L, G.inv_g);
        rs_R = (p_star > p_R) ? rho_R * (p_star + zeta * p_R) / (p_R + zeta * p_star)
                              : rho_R * gpow(p_star / p_R, G.inv_g);
        cs_L = sqrt(gamma * p_star / rs_L);
        cs_R = sqrt(gamma * p_star / rs_R);
    }

    if (dist < atc) {                                                    // acoustic (:79-139)
        if (u_L - c_L > 0.0) { edir_upwind(D, U, rho_L, u_L, p_L, s_rho_L, s_u_L, s_p_L, G); return; }
        if (u_R + c_R < 0.0) { edir_upwind(D, U, rho_R, u_R, p_R, s_rho_R, s_u_R, s_p_R, G); return; }
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

    const double speed_L = crw0 ? u_L - c_L : (rs_L * u_star - rho_L * u_L) / (rs_L - rho_L);
    const double speed_R = crw1 ? u_R + c_R : (rs_R * u_star - rho_R * u_R) / (rs_R - rho_R);

    if (speed_L > 0.0) {                                                 // left state on the axis (:158-171)
        D[0] = -s_rho_L * u_L - rho_L * s_u_L;
        D[1] = (D[0] * u_L + s_rho_L * u_L * u_L + 2.0 * rho_L * u_L * s_u_L + s_p_L) / -rho_L;
        D[2] = (s_u_L * p_L + u_L * s_p_L) * gamma / (1.0 - gamma) - 0.5 * s_rho_L * u_L * u_L * u_L
               - 1.5 * rho_L * u_L * u_L * s_u_L;
        D[2] = D[2] - 0.5 * D[0] * u_L * u_L - rho_L * u_L * D[1];
        D[2] = D[2] * (gamma - 1.0);
        U[0] = rho_L; U[1] = u_L; U[2] = p_L;
        return;
    }
    if (speed_R < 0.0) { edir_upwind(D, U, rho_R, u_R, p_R, s_rho_R, s_u_R, s_p_R, G); return; }

    if (crw0 && ((u_star - cs_L) > 0.0)) {                              // axis inside the 1-CRW (:188-205)
        U[1] = zeta * (u_L + 2.0 * c_L / (gamma - 1.0));
        U[2] = U[1] * U[1] * rho_L / gamma / gpow(p_L, G.inv_g);
        U[2] = gpow(U[2], G.g_gm1);
        U[0] = gamma * U[2] / U[1] / U[1];
        const PowBase b = pow_base(U[1] / c_L);
        const double pw_a = pow_with(b, G.e_a), pw_b = pow_with(b, G.e_b);
        D[1] = crw_coeff(pw_a, pw_b, zeta);
        D[1] = D[1] * (s_p_L - s_rho_L * c_L * c_L) / (gamma - 1.0) / rho_L;
        D[1] = D[1] - c_L * pw_a * (s_u_L + (gamma * s_p_L / c_L - c_L * s_rho_L) / (gamma - 1.0) / rho_L);
        D[2] = U[0] * U[1] * D[1];
        D[0] = U[0] * U[1] * pw_b * (s_p_L - s_rho_L * c_L * c_L) / rho_L;
        D[0] = (D[0] + D[2]) / U[1] / U[1];
        return;
    }
    if (crw1 && ((u_star + cs_R) < 0.0)) {                              // axis inside the 3-CRW (:206-223)
        U[1] = zeta * (u_R - 2.0 * c_R / (gamma - 1.0));
        U[2] = U[1] * U[1] * rho_R / gamma / gpow(p_R, G.inv_g);
        U[2] = gpow(U[2], G.g_gm1);
        U[0] = gamma * U[2] / U[1] / U[1];
        const PowBase b = pow_base(-U[1] / c_R);
        const double pw_a = pow_with(b, G.e_a), pw_b = pow_with(b, G.e_b);
        D[1] = crw_coeff(pw_a, pw_b, zeta);
        D[1] = D[1] * (s_p_R - s_rho_R * c_R * c_R) / (gamma - 1.0) / rho_R;
        D[1] = D[1] + c_R * pw_a * (s_u_R - (gamma * s_p_R / c_R - c_R * s_rho_R) / (gamma - 1.0) / rho_R);
        D[2] = U[0] * U[1] * D[1];
        D[0] = U[0] * U[1] * pw_b * (s_p_R - s_rho_R * c_R * c_R) / rho_R;
        D[0] = (D[0] + D[2]) / U[1] / U[1];
        return;
    }

    double a_L, b_L, d_L, a_R, b_R, d_R, pwL_b = 0.0, pwR_b = 0.0;      // non-sonic (:225-327)
    if (crw0) {
        a_L = 1.0;
        b_L = 1.0 / rs_L / cs_L;
        const PowBase b = pow_base(cs_L / c_L);
        const double pw_a = pow_with(b, G.e_a);
        pwL_b = pow_with(b, G.e_b);
        d_L = crw_coeff(pw_a, pwL_b, zeta);
        d_L = d_L * (s_p_L - s_rho_L * c_L * c_L) / (gamma - 1.0) / rho_L;
        d_L = d_L - c_L * pw_a * (s_u_L + (gamma * s_p_L / c_L - c_L * s_rho_L) / (gamma - 1.0) / rho_L);
    } else {
        const double VAR = sqrt((1.0 - zeta) / (rho_L * (p_star + zeta * p_L)));
        const double H1 = 0.5 * VAR * (p_star + (1.0 + 2.0 * zeta) * p_L) / (p_star + zeta * p_L);
        const double H2 = -0.5 * VAR * ((2.0 + zeta) * p_star + zeta * p_L) / (p_star + zeta * p_L);
        const double H3 = -0.5 * VAR * (p_star - p_L) / rho_L;
        const double shk = (rs_L * u_star - rho_L * u_L) / (rs_L - rho_L);
        a_L = 1.0 - rs_L * (shk - u_star) * H1;
        b_L = (u_star - shk) / rs_L / cs_L / cs_L + H1;
        const double L_rho = (u_L - shk) * H3;
        const double L_u = shk - u_L + rho_L * c_L * c_L * H2 + rho_L * H3;
        const double L_p = (u_L - shk) * H2 - 1.0 / rho_L;
        d_L = L_rho * s_rho_L + L_u * s_u_L + L_p * s_p_L;
    }
    if (crw1) {
        a_R = 1.0;
        b_R = -1.0 / rs_R / cs_R;
        const PowBase b = pow_base(cs_R / c_R);
        const double pw_a = pow_with(b, G.e_a);
        pwR_b = pow_with(b, G.e_b);
        d_R = crw_coeff(pw_a, pwR_b, zeta);
        d_R = d_R * (s_p_R - s_rho_R * c_R * c_R) / (gamma - 1.0) / rho_R;
        d_R = d_R + c_R * pw_a * (s_u_R - (gamma * s_p_R / c_R - c_R * s_rho_R) / (gamma - 1.0) / rho_R);
    } else {
        const double VAR = sqrt((1.0 - zeta) / (rho_R * (p_star + zeta * p_R)));
        const double H1 = 0.5 * VAR * (p_star + (1.0 + 2.0 * zeta) * p_R) / (p_star + zeta * p_R);
        const double H2 = -0.5 * VAR * ((2.0 + zeta) * p_star + zeta * p_R) / (p_star + zeta * p_R);
        const double H3 = -0.5 * VAR * (p_star - p_R) / rho_R;
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
        if (crw1) {
            D[0] = rs_R * u_star * pwR_b * (s_p_R - s_rho_R * c_R * c_R) / rho_R;
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
        if (crw0) {
            D[0] = rs_L * u_star * pwL_b * (s_p_L - s_rho_L * c_L * c_L) / rho_L;
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

// the 1D solver's call: atc = eps (grp_solver_EUL_source.c)
__device__ inline void grp_edir(double D[3], double U[3], const Side1 &L, const Side1 &R, const Gas &G, double eps)
{
    grp_edir(D, U, L, R, G, eps, eps);
}

}  // namespace hgrp
