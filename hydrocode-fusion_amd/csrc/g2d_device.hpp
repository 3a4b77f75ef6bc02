// g2d_device.hpp -- the genuinely-2D direct Eulerian GRP solver on gfx950 (FP64).
//
//   grp_g2d <- linear_GRP_solver_Edir_G2D  src/riemann_solver/linear_grp_solver_Edir_G2D.c:48-637
//
// The reference's default build (EXACT_TANGENT_DERIVATIVE undefined): the
// tangential derivatives on the t-axis follow from the acoustic relations
// across the star state of the contact's side (:337-349) and enter the
// characteristic relations of both waves, so the time derivatives are those
// of the genuinely two-dimensional problem rather than Q1D's quasi-1D ones
// plus transversal terms.  Unlike Q1D the solver's distance between the sides
// includes v, lambda_v is the right side's (:50), and a gas at gamma = 3 takes
// the log limit of the rarefaction's tangential term.
//
// One code path serves both waves with the wave's direction sigma (-1: the
// 1-wave on the left, +1: the 3-wave on the right); each mirrored expression
// of the reference is one IEEE arithmetic evaluates to the same double
// (x - y and x + (-y), (-a) b and -(a b), -(a + b) and (-a) - b), and where a
// mirror would change the sign of a zero (the sonic point's C) the two sides
// are written out.  Expression order is the reference's throughout, so with
// -ffp-contract=off, IEEE division and square root and the accurate pow of
// fastpow.hpp the results are the reference's to the last bit except where
// glibc's pow or log is not correctly rounded.
#pragma once
#include "grp_device.hpp"

namespace hgrp {

// one side of the interface with the members the solver reads
struct G2Side {
    double sg;                                   // -1 left (1-wave), +1 right (3-wave)
    Gas G;
    double c, rho, u, v, p;
    double d_rho, d_u, d_v, d_p;
    double t_rho, t_u, t_v, t_p;
    double z, d_z, t_z, phi, d_phi, t_phi;       // multi-fluid members (constants without them)
    double rs, cs;                               // star density and sound speed on this side
    int crw;                                     // this side's wave is a rarefaction
};

// The interface sides of the C-ABI: 21 doubles = gamma, lambda_u, lambda_v,
// rho, u, v, p, d_rho, d_u, d_v, d_p, t_rho, t_u, t_v, t_p, z_a, d_z_a, t_z_a,
// phi, d_phi, t_phi.
__device__ __forceinline__ G2Side g2_side_of(const double *a, double sg)
{
    G2Side s;
    s.sg = sg;
    s.G = make_gas(a[0]);
    s.rho = a[3]; s.u = a[4]; s.v = a[5]; s.p = a[6];
    s.d_rho = a[7]; s.d_u = a[8]; s.d_v = a[9]; s.d_p = a[10];
    s.t_rho = a[11]; s.t_u = a[12]; s.t_v = a[13]; s.t_p = a[14];
    s.z = a[15]; s.d_z = a[16]; s.t_z = a[17]; s.phi = a[18]; s.d_phi = a[19]; s.t_phi = a[20];
    s.c = sqrt(s.G.g * s.p / s.rho);
    s.crw = 0;
    return s;
}

// star density and sound speed of the side, its outer wave speed (:118-150)
__device__ __forceinline__ double g2_star_side(G2Side &s, double p_star)
{
    const Gas &G = s.G;
    if (s.crw) {
        const PowBase b = pow_base(p_star / s.p);
        s.rs = s.rho * pow_with(b, G.inv_g);
        s.cs = s.c * pow_with(b, G.e_cs);
        return s.u + s.sg * s.c;
    }
    s.rs = s.rho * (p_star + G.z * s.p) / (s.p + G.z * p_star);
    s.cs = sqrt(G.g * p_star / s.rs);
    return s.u + s.sg * (s.c * sqrt(0.5 * (G.gp1() * (p_star / s.p) + G.gm1()) / G.g));
}

// the t-axis outside the wave fan: upwind characteristic derivative
__device__ __forceinline__ void g2_upwind(double D[6], double U[6], const G2Side &s, double lu, double lv)
{
    const double a = s.u - lu, b = s.v - lv;
    U[0] = s.rho; U[1] = s.u; U[2] = s.v; U[3] = s.p; U[4] = s.z; U[5] = s.phi;
    D[0] = -a * s.d_rho - b * s.t_rho - s.rho * (s.d_u + s.t_v);
    D[1] = -a * s.d_u   - b * s.t_u   - s.d_p / s.rho;
    D[2] = -a * s.d_v   - b * s.t_v   - s.t_p / s.rho;
    D[3] = -a * s.d_p   - b * s.t_p   - s.rho * s.c * s.c * (s.d_u + s.t_v);
    D[4] = -a * s.d_z   - b * s.t_z;
    D[5] = -a * s.d_phi - b * s.t_phi;
}

// the sonic point of a rarefaction fan containing the t-axis; returns C there
// and the PowBase of c_frac = C / c for the powers that follow
__device__ __forceinline__ double g2_sonic(double U[6], const G2Side &s, double lu, PowBase &cf)
{
    const Gas &G = s.G;
    double C;
    if (s.sg < 0.0) {
        U[1] = G.z * (s.u + 2.0 * (s.c + lu) / G.gm1());
        C = U[1] - lu;
    } else {
        U[1] = G.z * (s.u - 2.0 * (s.c - lu) / G.gm1());
        C = lu - U[1];
    }
    cf = pow_base(C / s.c);
    U[3] = pow_with(cf, G.e_sonic) * s.p;
    U[0] = G.g * U[3] / C / C;
    U[2] = s.v;
    U[4] = s.z;
    U[5] = s.phi;
    return C;
}

// the rarefaction's tangential coefficient Q (:368-371); gamma = 3 has its own limit
__device__ __forceinline__ double g2_q(const Gas &G, double c_frac, double pw_a, double eps)
{
    const double z = G.z;
    if (G.g < 3.0 - eps || G.g > 3.0 + eps) return (c_frac * (z - 1.0) + pw_a * z) / (2.0 * z - 1.0);
    return 0.5 * c_frac + pw_a * (0.5 - 0.25 / z * log(c_frac));
}

// right-hand side of the rarefaction's characteristic relation at c_frac
// (:362-372, :431-441); pw_a = c_frac^(0.5/zeta), pw_b = c_frac^((1+zeta)/zeta)
__device__ __forceinline__ double g2_crw_rhs(const G2Side &s, double c_frac, double pw_a, double pw_b, double eps)
{
    const Gas &G = s.G;
    const double TdS = (s.d_p - s.d_rho * s.c * s.c) / G.gm1() / s.rho;
    const double d_R = s.d_u - s.sg * ((G.g * s.d_p / s.c - s.c * s.d_rho) / G.gm1() / s.rho);
    double d = (G.opz() * pw_a + G.z * pw_b);
    d = d / G.op2z() * TdS;
    d = d + s.sg * (s.c * pw_a * d_R);
    return d + s.sg * (s.c * s.t_v * g2_q(G, c_frac, pw_a, eps));
}

// the t-axis inside a rarefaction fan (:351-428)
__device__ __forceinline__ void g2_in_fan(double D[6], double U[6], const G2Side &s, double lu, double lv,
                                          double T_rho, double T_u, double T_p, double eps)
{
    const Gas &G = s.G;
    PowBase cf;
    const double C = g2_sonic(U, s, lu, cf);
    const double z = G.z, c_frac = C / s.c;
    const double pw_a = pow_with(cf, G.e_a), pw_b = pow_with(cf, G.e_b);
    const double TdS = (s.d_p - s.d_rho * s.c * s.c) / G.gm1() / s.rho;
    const double w = U[1] - lu, b = U[2] - lv;
    D[1] = g2_crw_rhs(s, c_frac, pw_a, pw_b, eps);
    D[3] = U[0] * w * D[1];
    D[0] = U[0] * w * pw_b * TdS * G.gm1();
    D[0] = (D[0] + D[3]) / C / C - b * T_rho;
    D[2] = -w * s.d_v * U[0] / s.rho - b * s.t_v - T_p / U[0];
    D[2] = D[2] - (z - 1.0) * (pow_with(cf, 2.0 / z - 1.0) - 1.0) / (z - 2.0) / U[0] * s.t_p;
    D[4] = -w * s.d_z * U[0] / s.rho - b * s.t_z;
    D[5] = -w * s.d_phi * U[0] / s.rho - b * s.t_phi;
    D[3] = D[3] - b * T_p;
    D[1] = D[1] - b * T_u + U[1] * s.t_v;
}

// shock speeds relative to the star state and to the unshocked side, signed
// with the wave (:443-444, :472-473)
__device__ __forceinline__ void g2_shock_speeds(const G2Side &s, double p_star, double &SmUs, double &SmU)
{
    const Gas &G = s.G;
    SmUs = s.sg * sqrt(0.5 * (G.gp1() * s.p + G.gm1() * p_star) / s.rs);
    SmU = s.sg * sqrt(0.5 * (G.gp1() * p_star + G.gm1() * s.p) / s.rho);
}

// one row a u_t + b p_t = d of the 2x2 system on the t-axis (:429-491)
__device__ __forceinline__ void g2_wave_row(const G2Side &s, double p_star, double T_u, double T_p, double eps,
                                            double &a, double &b, double &d)
{
    const Gas &G = s.G;
    if (s.crw) {
        a = 1.0;
        b = -s.sg * (1.0 / s.rs / s.cs);
        const PowBase cf = pow_base(s.cs / s.c);
        d = g2_crw_rhs(s, s.cs / s.c, pow_with(cf, G.e_a), pow_with(cf, G.e_b), eps);
    } else {
        double SmUs, SmU;
        g2_shock_speeds(s, p_star, SmUs, SmU);
        const double z = G.z, pz = p_star + z * s.p;
        const double VAR = sqrt(G.omz() / (s.rho * pz));
        const double H1 = 0.5 * VAR * (p_star + G.op2z() * s.p) / pz;
        const double H2 = -0.5 * VAR * (G.tpz() * p_star + z * s.p) / pz;
        const double H3 = -0.5 * VAR * (p_star - s.p) / s.rho;
        const double k = s.rho * (s.c * s.c * H2 + H3);
        const double L_p = -1.0 / s.rho + s.sg * (SmU * H2);
        const double L_u = SmU - s.sg * k;
        const double L_rho = s.sg * (SmU * H3);
        const double L_v = SmUs - s.sg * k;
        a = 1.0 + s.sg * (s.rs * SmUs * H1);
        b = -(SmUs / (s.rs * s.cs * s.cs)) - s.sg * H1;
        d = L_rho * s.d_rho + L_u * s.d_u + L_p * s.d_p + L_v * s.t_v;
    }
    d = d - a * s.v * T_u - b * s.v * T_p;
}

// linear_GRP_solver_Edir_G2D.  lu: lambda_u of the left side, lv: lambda_v of
// the right side.  D[6], U[6] = rho u v p z_a phi; ws[2] wave speeds; us[6] =
// rho_star_L u_star rho_star_R p_star c_star_L c_star_R.  tol, p_guess and
// CONV as riemann_exact's (the reference: tol = eps, no guess, its own rule).
template <bool CONV = false>
__device__ __forceinline__ void grp_g2d(double D[6], double U[6], double ws[2], double us[6], G2Side L, G2Side R,
                                        double lu, double lv, double eps, double atc, double tol,
                                        double p_guess = -1.0)
{
    double u_star, p_star, speed_L, speed_R;
    const double dist = sqrt((L.rho - R.rho) * (L.rho - R.rho) + (L.u - R.u) * (L.u - R.u)
                             + (L.v - R.v) * (L.v - R.v) + (L.p - R.p) * (L.p - R.p));
    if (dist < atc && atc < 2 * eps) {
        u_star = 0.5 * (R.u + L.u);
        p_star = 0.5 * (R.p + L.p);
        L.rs = L.rho; L.cs = L.c; speed_L = u_star - L.cs;
        R.rs = R.rho; R.cs = R.c; speed_R = u_star + R.cs;
    } else {
        const StarResult st = riemann_exact<CONV>(L.G, R.G, L.u, R.u, L.p, R.p, L.c, R.c, eps, tol, 500, p_guess);
        u_star = st.u_star;
        p_star = st.p_star;
        L.crw = st.crw0;
        R.crw = st.crw1;
        speed_L = g2_star_side(L, p_star);
        speed_R = g2_star_side(R, p_star);
    }
    ws[0] = speed_L; ws[1] = speed_R;
    us[0] = L.rs; us[1] = u_star; us[2] = R.rs; us[3] = p_star; us[4] = L.cs; us[5] = R.cs;

    if (speed_L > lu) { g2_upwind(D, U, L, lu, lv); return; }
    if (speed_R < lu) { g2_upwind(D, U, R, lu, lv); return; }
    const bool fan_L = L.crw && ((u_star - L.cs) > lu);
    const bool fan_R = R.crw && ((u_star + R.cs) < lu);

    if (dist < atc) {
        // acoustic case (:157-254): the side of the contact decides v, z, phi
        // and the entropy term
        const G2Side &S = u_star > lu ? L : R;
        double C;
        PowBase cf;
        if (fan_L) C = g2_sonic(U, L, lu, cf);
        else if (fan_R) C = g2_sonic(U, R, lu, cf);
        else {
            U[0] = S.rs; U[1] = u_star; U[2] = S.v; U[3] = p_star; U[4] = S.z; U[5] = S.phi;
            C = S.cs;
        }
        const double D_p = 0.5 * ((L.d_u * (U[0] * C) + L.d_p) - (R.d_u * (U[0] * C) - R.d_p));
        const double T_p = 0.5 * ((L.t_u * (U[0] * C) + L.t_p) - (R.t_u * (U[0] * C) - R.t_p));
        const double D_u = 0.5 * (L.d_u + L.d_p / (U[0] * C) + R.d_u - R.d_p / (U[0] * C));
        const double T_u = 0.5 * (L.t_u + L.t_p / (U[0] * C) + R.t_u - R.t_p / (U[0] * C));
        const double D_rho = S.d_rho - S.d_p / (C * C) + D_p / (C * C);
        const double T_rho = S.t_rho - S.t_p / (C * C) + T_p / (C * C);
        const double w = U[1] - lu, b = U[2] - lv;
        D[0] = -w * D_rho   - b * T_rho   - U[0] * (D_u + S.t_v);
        D[1] = -w * D_u     - b * T_u     - D_p / U[0];
        D[2] = -w * S.d_v   - b * S.t_v   - T_p / U[0];
        D[3] = -w * D_p     - b * T_p     - U[0] * C * C * (D_u + S.t_v);
        D[4] = -w * S.d_z   - b * S.t_z;
        D[5] = -w * S.d_phi - b * S.t_phi;
        return;
    }

    // tangential derivatives on the t-axis (:337-349)
    const G2Side &T = u_star < lu ? R : L;
    const double T_p = 0.5 * ((L.t_u - R.t_u) * T.rs * T.cs + L.t_p + R.t_p);
    const double T_u = 0.5 * (L.t_u + R.t_u + (L.t_p - R.t_p) / T.rs / T.cs);
    const double T_rho = T.t_rho - T.t_p / (T.cs * T.cs) + T_p / (T.cs * T.cs);

    if (fan_L) { g2_in_fan(D, U, L, lu, lv, T_rho, T_u, T_p, eps); return; }
    if (fan_R) { g2_in_fan(D, U, R, lu, lv, T_rho, T_u, T_p, eps); return; }

    // non-sonic (:429-627)
    U[0] = T.rs; U[1] = u_star; U[2] = T.v; U[3] = p_star; U[4] = T.z; U[5] = T.phi;
    const double C = T.cs;
    double a_L, b_L, d_L, a_R, b_R, d_R;
    g2_wave_row(L, p_star, T_u, T_p, eps, a_L, b_L, d_L);
    g2_wave_row(R, p_star, T_u, T_p, eps, a_R, b_R, d_R);
    const double detA = a_L * b_R - b_L * a_R;
    const double u_t_mat = (b_R * d_L - b_L * d_R) / detA;
    const double p_t_mat = (a_L * d_R - a_R * d_L) / detA;
    const double D0_p_tau = p_t_mat + U[2] * T_p;
    const double D0_u_tau = u_t_mat + U[2] * T_u;
    const double w = u_star - lu, b = U[2] - lv;
    D[1] = u_t_mat + w / U[0] / C / C * D0_p_tau + w * T.t_v;
    D[3] = p_t_mat + w * U[0] * D0_u_tau;
    const Gas &G = T.G;
    if (T.crw) {
        // the reference's c_frac here is the last one it assigned: the right
        // wave's when that is a rarefaction, else the left wave's (:432, :462)
        const double c_frac = R.crw ? R.cs / R.c : L.cs / L.c;
        const double z = G.z;
        D[0] = T.rs * w * pow_with(pow_base(T.cs / T.c), G.e_b) * (T.d_p - T.d_rho * T.c * T.c) / T.rho;
        D[0] = (D[0] + D[3] + U[2] * T_p) / T.cs / T.cs - b * T_rho;
        D[2] = -U[1] * T.d_v * U[0] / T.rho - b * T.t_v - T_p / U[0];
        D[2] = D[2] + lu * T.d_v;
        D[2] = D[2] + T.sg * (u_star / T.cs * (z - 1.0) * (gpow(c_frac, 2.0 / z - 1.0) - 1.0) / (z - 2.0) / U[0]
                              * T.t_p);
        D[4] = -U[1] * T.d_z * U[0] / T.rho - b * T.t_z;
        D[4] = D[4] + lu * T.d_z;
        D[5] = -U[1] * T.d_phi * U[0] / T.rho - b * T.t_phi;
        D[5] = D[5] + lu * T.d_phi;
    } else {
        double SmUs, SmU;
        g2_shock_speeds(T, p_star, SmUs, SmU);
        const double z = G.z, r = T.rho, c = T.c;
        const double VAR = T.p + z * p_star;
        const double H1 = r * T.p * G.omzz() / VAR / VAR;
        const double H2 = r * p_star * G.zzm1() / VAR / VAR;
        const double H3 = (p_star + z * T.p) / VAR;
        const double L_rho = SmU * H3 * T.d_rho;
        const double L_u = -r * (H2 * c * c + H3) * T.d_u;
        const double L_p = H2 * SmU * T.d_p;
        const double L_v = -r * (H2 * c * c + H3) * T.t_v;
        D[0] = ((u_star + SmUs) / T.cs / T.cs - u_star * H1) * D0_p_tau + T.rs * u_star * SmUs * H1 * D0_u_tau;
        D[0] = (D[0] - u_star * (L_p + L_rho + L_u + L_v)) / SmUs;
        D[0] = D[0] - b * T_rho;
        const double f = SmU * (H2 * T.d_p + H3 * T.d_rho) - r * (H2 * c * c + H3) * T.d_u;
        const double rho_x = (f + H1 * (p_t_mat - T.rs * SmUs * u_t_mat) - D[0]) / (SmU + T.u);
        D[0] = D[0] + lu * rho_x;
        D[2] = -(U[1] * (SmU * T.d_v - T.t_p / r) + (u_star + SmUs) * T_p / U[0]) / SmUs;
        D[2] = D[2] + lu * T.d_v - b * T.t_v;
        D[4] = -U[1] * SmU * T.d_z / SmUs;
        D[4] = D[4] + lu * T.d_z - b * T.t_z;
        D[5] = -U[1] * SmU * T.d_phi / SmUs;
        D[5] = D[5] + lu * T.d_phi - b * T.t_phi;
    }
    D[1] = D[1] + lv * T_u;
    D[3] = D[3] + lv * T_p;
}

// An interface side of the 2D step (single fluid: z_a and phi take the
// reference's single-fluid constants 0, -0).  T: tangential derivatives
// (rho, u, v, p) in the interface's frame.
__device__ __forceinline__ G2Side g2_side(const Side &S, const double T[4], const Gas &G, double sg)
{
    G2Side s;
    s.sg = sg;
    s.G = G;
    s.rho = S.rho; s.u = S.u; s.v = S.v; s.p = S.p;
    s.d_rho = S.d_rho; s.d_u = S.d_u; s.d_v = S.d_v; s.d_p = S.d_p;
    s.t_rho = T[0]; s.t_u = T[1]; s.t_v = T[2]; s.t_p = T[3];
    s.z = 0.0; s.d_z = -0.0; s.t_z = -0.0; s.phi = 0.0; s.d_phi = -0.0; s.t_phi = -0.0;
    s.c = sqrt(G.g * s.p / s.rho);
    s.crw = 0;
    return s;
}

// GRP_2D_flux with the reference's alternative interface solver, G2D at
// atc = eps (src/flux_calc/flux_solver.c:189 in place of the Q1D call at
// :191), in the unsplit step's form: one gas, normal (n_x, n_y), both sides'
// tangential derivatives TL, TR in grid axes.  GRP_2D_flux zeroes the left
// side's grid velocity and the flux generators leave the right side's at 0, so
// the solver sees lambda_u = lambda_v = 0.  Results and codes as grp_2d_flux_g.
template <bool CONV = false>
__device__ __forceinline__ int grp_2d_flux_g2d(Side L, Side R, const double TL[4], const double TR[4], const Gas &G,
                                               double n_x, double n_y, double tau, double eps, double F[4],
                                               double I[4], double tol, double p_guess = -1.0)
{
    const double u = L.u * n_x + L.v * n_y;
    const double u_R = R.u * n_x + R.v * n_y;
    const double d_u = L.d_u * n_x + L.d_v * n_y;
    const double d_u_R = R.d_u * n_x + R.d_v * n_y;
    const double TLr[4] = {TL[0], TL[1] * n_x + TL[2] * n_y, -TL[1] * n_y + TL[2] * n_x, TL[3]};
    const double TRr[4] = {TR[0], TR[1] * n_x + TR[2] * n_y, -TR[1] * n_y + TR[2] * n_x, TR[3]};
    L.v = -L.u * n_y + L.v * n_x;
    R.v = -R.u * n_y + R.v * n_x;
    L.d_v = -L.d_u * n_y + L.d_v * n_x;
    R.d_v = -R.d_u * n_y + R.d_v * n_x;
    L.u = u; R.u = u_R; L.d_u = d_u; R.d_u = d_u_R;
    double dire[6], mid[6], ws[2], us[6];
    grp_g2d<CONV>(dire, mid, ws, us, g2_side(L, TLr, G, -1.0), g2_side(R, TRr, G, 1.0), 0.0, 0.0, eps, eps, tol,
                  p_guess);
    return flux_from_mid_n(mid, dire, G, n_x, n_y, tau, eps, F, I);
}

}  // namespace hgrp
