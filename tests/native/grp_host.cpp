// Host build of the device GRP solvers (hydrocode-fusion_amd/csrc/grp_device.hpp,
// grp1d_device.hpp, g2d_device.hpp) so that their arithmetic can be compared with the oracle
// on the CPU: same source, same expression order, no FMA contraction.
#include <hip/hip_runtime.h>
#include "grp_device.hpp"
#include "grp1d_device.hpp"
#include "g2d_device.hpp"

using namespace hgrp;

// L, R: gamma lambda_u lambda_v rho u v p d_rho d_u d_v d_p ... (orc_ifv order, 15 doubles)
extern "C" void hg_edir_many(const double *L, const double *R, int n, double eps, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side1 a{l[3], l[4], l[6], l[7], l[8], l[10]}, b{r[3], r[4], r[6], r[7], r[8], r[10]};
        double D[3], U[3];
        grp_edir(D, U, a, b, make_gas(l[0]), eps);
        for (int q = 0; q < 3; ++q) { out[6 * k + q] = D[q]; out[6 * k + 3 + q] = U[q]; }
    }
}

extern "C" void hg_flux_many(const double *L, const double *R, int n, double tau, double eps, int dir, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side a{l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]}, b{r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
        double F[4], I[4];
        const int rc = dir == 0 ? grp_2d_flux<0>(a, b, make_gas(l[0]), tau, eps, F, I, eps)
                                : grp_2d_flux<1>(a, b, make_gas(l[0]), tau, eps, F, I, eps);
        out[9 * k] = rc;
        for (int q = 0; q < 4; ++q) { out[9 * k + 1 + q] = F[q]; out[9 * k + 5 + q] = I[q]; }
    }
}

// the unsplit path's flux: tangential derivatives of both sides (members 11..14), run-time normal
extern "C" void hg_flux_t_many(const double *L, const double *R, int n, double tau, double eps, int dir, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side a{l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]}, b{r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
        double F[4], I[4];
        const int rc = grp_2d_flux_n(a, b, l + 11, r + 11, make_gas(l[0]), dir ? 0.0 : 1.0, dir ? 1.0 : 0.0, tau,
                                     eps, F, I, eps);
        out[9 * k] = rc;
        for (int q = 0; q < 4; ++q) { out[9 * k + 1 + q] = F[q]; out[9 * k + 5 + q] = I[q]; }
    }
}

extern "C" void hg_div_by_many(const double *a, int n, double b, double *out)
{
    const RecipDiv d = recip_div(b);
    for (int k = 0; k < n; ++k) out[k] = div_by(a[k], d);
}

// the per-interface entry points' forms: atc, grid velocity of L, a gas per side
extern "C" void hg_q1d_many(const double *L, const double *R, int n, double eps, double atc, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side a{l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]}, b{r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
        double D[4], U[4], ws[2], us[6];
        const Q1DExtra X = {atc, l[1], l[2], ws, us};
        grp_q1d(D, U, a, b, l + 11, r + 11, make_gas(l[0]), make_gas(r[0]), eps, eps, -1.0, nullptr, nullptr, &X);
        double *o = out + 16 * k;
        o[0] = ws[0]; o[1] = ws[1];
        for (int q = 0; q < 4; ++q) { o[2 + q] = D[q]; o[6 + q] = U[q]; }
        for (int q = 0; q < 6; ++q) o[10 + q] = us[q];
    }
}

// linear_GRP_solver_Edir_G2D on sides of 21 doubles; out: ws[2] D[6] U[6] U_star[6]
extern "C" void hg_g2d_many(const double *L, const double *R, int n, double eps, double atc, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 21 * k, *r = R + 21 * k;
        double D[6], U[6], ws[2], us[6];
        grp_g2d(D, U, ws, us, g2_side_of(l, -1.0), g2_side_of(r, 1.0), l[1], r[2], eps, atc, eps);
        double *o = out + 20 * k;
        o[0] = ws[0]; o[1] = ws[1];
        for (int q = 0; q < 6; ++q) { o[2 + q] = D[q]; o[8 + q] = U[q]; o[14 + q] = us[q]; }
    }
}

extern "C" void hg_flux_g_many(const double *L, const double *R, int n, double tau, double eps, double n_x,
                               double n_y, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side a{l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]}, b{r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
        double F[4], I[4];
        const int rc = grp_2d_flux_g(a, b, l + 11, r + 11, make_gas(l[0]), make_gas(r[0]), n_x, n_y, tau, eps, F, I,
                                     eps);
        out[9 * k] = rc;
        for (int q = 0; q < 4; ++q) { out[9 * k + 1 + q] = F[q]; out[9 * k + 5 + q] = I[q]; }
    }
}

extern "C" void hg_edir_atc_many(const double *L, const double *R, int n, double eps, double atc, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *l = L + 15 * k, *r = R + 15 * k;
        Side1 a{l[3], l[4], l[6], l[7], l[8], l[10]}, b{r[3], r[4], r[6], r[7], r[8], r[10]};
        double D[3], U[3];
        grp_edir(D, U, a, b, make_gas(l[0]), eps, atc);
        for (int q = 0; q < 3; ++q) { out[6 * k + q] = D[q]; out[6 * k + 3 + q] = U[q]; }
    }
}

// in: gammaL gammaR u_L u_R p_L p_R c_L c_R eps tol; out: U* P* CRW0 CRW1 gap iters
extern "C" void hg_riemann_many(const double *in, int n, int N, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *a = in + 10 * k;
        const StarResult r = riemann_exact(make_gas(a[0]), make_gas(a[1]), a[2], a[3], a[4], a[5], a[6], a[7], a[8],
                                           a[9], N);
        double *o = out + 6 * k;
        o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.gap; o[5] = r.iters;
    }
}

// GuessP per row: DL DR UL UR PL PR CL CR GammaL GammaR eps
extern "C" void hg_guessp_many(const double *in, int n, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *a = in + 11 * k;
        out[k] = guess_pvrs(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);
    }
}

// riemann_exact from a given starting pressure (guess[k] <= 0: the reference's guess)
extern "C" void hg_riemann_guess_many(const double *in, int n, int N, const double *guess, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *a = in + 10 * k;
        const StarResult r = riemann_exact(make_gas(a[0]), make_gas(a[1]), a[2], a[3], a[4], a[5], a[6], a[7], a[8],
                                           a[9], N, guess[k]);
        double *o = out + 6 * k;
        o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.gap; o[5] = r.iters;
    }
}

// riemann_exact with a stopping rule (0 the reference's, 1 the converge rule)
// from a given starting pressure (guess[k] <= 0: the reference's guess)
extern "C" void hg_riemann_rule_many(const double *in, int n, int N, int rule, const double *guess, double *out)
{
    for (int k = 0; k < n; ++k) {
        const double *a = in + 10 * k;
        const Gas GL = make_gas(a[0]), GR = make_gas(a[1]);
        const StarResult r =
            rule ? riemann_exact<true>(GL, GR, a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], N, guess[k])
                 : riemann_exact<false>(GL, GR, a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], N, guess[k]);
        double *o = out + 6 * k;
        o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.gap; o[5] = r.iters;
    }
}

// GRP_2D_flux with the two-fluid members, as flux_tf_batch_kernel (21-double sides)
extern "C" void hg_flux_tf_many(const double *L, const double *R, int n, double tau, double eps, int dir, double ga,
                                double gb, double *out)
{
    TwoGas T;
    T.a = make_gas(ga);
    T.b = make_gas(gb);
    for (int k = 0; k < n; ++k) {
        const double *l = L + 21 * k, *r = R + 21 * k;
        Side SL = {l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]};
        Side SR = {r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
        const double FA[4] = {l[15], l[16], r[15], r[16]};
        double F[4], I[4], Z[3];
        const int code = dir == 0 ? grp_2d_flux_tf<0>(SL, SR, T, FA, tau, eps, F, I, Z, eps)
                                  : grp_2d_flux_tf<1>(SL, SR, T, FA, tau, eps, F, I, Z, eps);
        double *o = out + 10 * k;
        o[0] = code;
        for (int q = 0; q < 4; ++q) { o[1 + q] = F[q]; o[5 + q] = I[q]; }
        o[9] = Z[2];
    }
}
