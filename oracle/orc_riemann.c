/*
 * orc_riemann.c -- exact Riemann solvers of the reference, restated.
 * TEST INFRASTRUCTURE ONLY (the oracle); see grp_oracle.h.
 *
 * Follows src/riemann_solver/riemann_solver_exact_Ben.c:
 *   orc_riemann_exact     <- Riemann_solver_exact      (:30-211)
 *   orc_riemann_exact_ben <- Riemann_solver_exact_Ben  (:230-408)
 * and src/riemann_solver/riemann_solver_starPU.c:
 *   orc_guessp            <- GuessP                    (:20-58)
 * Both evaluate the two wave curves through the left/right states
 * (shock branch above the state pressure, isentrope below), classify the two
 * waves with an eps margin, start Newton from half the intersection of the two
 * acoustic tangents and iterate on the absolute velocity gap.
 *
 * orc_riemann_exact_rule adds the converge stopping rule (no counterpart in the
 * reference): an iterate within tol that the starting value or an ordinary
 * step gave gets one more Newton step when none has been taken yet or when
 * its gap exceeds the smaller of 2^-49 (|v_L| + |v_R| + c_L + c_R), the
 * rounding floor of the curves' velocities, and 1e-13 p |k1 - k3| (p more than
 * 1e-13 off the root by the last tangents); an iterate that such an
 * extra step gave ends the loop if it is within tol.  Pinned by the root of the
 * wave-curve equation computed in 40-digit arithmetic (tests/test_oracle_math.py).
 */
#include <math.h>
#include "grp_oracle.h"

/* Velocity on the 1-wave curve through (u_s, p_s) at pressure p ("-" side) or
 * on the 3-wave curve ("+" side).  Expression order as at :53-61 / :99-107. */
static double wave_curve(double p, double u_s, double p_s, double c_s, double g,
                         double mu, double nu, double sign)
{
    double w;
    if (p > p_s) {
        double dp = p - p_s;
        double s = sqrt(1.0 + nu * dp / p_s);
        w = dp * c_s / g / p_s / s;
    } else {
        w = pow(p / p_s, mu) - 1.0;
        w = 2.0 * c_s * w / (g - 1.0);
    }
    return sign < 0.0 ? u_s - w : u_s + w;
}

static double riemann_core(double *U_star, double *P_star,
                           double gL, double gR, double muL, double nuL, double muR, double nuR,
                           double u_L, double u_R, double p_L, double p_R,
                           double c_L, double c_R, int CRW[2],
                           double eps, double tol, int N, int *n_iter, int conv)
{
    /* wave types (:50-84): (u_LR, p_R) on the left curve, (u_RL, p_L) on the right */
    const double u_LR = wave_curve(p_R, u_L, p_L, c_L, gL, muL, nuL, -1.0);
    const double u_RL = wave_curve(p_L, u_R, p_R, c_R, gR, muR, nuR, +1.0);
    CRW[1] = (u_LR > u_R + eps) ? 0 : 1;
    CRW[0] = (u_RL > u_L - eps) ? 1 : 0;

    /* halved intersection of the acoustic tangents (:87-93) */
    double k1 = -c_L / p_L / gL;
    double k3 = c_R / p_R / gR;
    double p_INT = (k1 * p_L - k3 * p_R - u_L + u_R) / (k1 - k3);
    if (p_INT < 0)
        p_INT = (p_L < p_R) ? p_L : p_R;
    p_INT = 0.5 * p_INT;

    double v_L = wave_curve(p_INT, u_L, p_L, c_L, gL, muL, nuL, -1.0);
    double v_R = wave_curve(p_INT, u_R, p_R, c_R, gR, muR, nuR, +1.0);
    double gap = fabs(v_L - v_R);

    if (fabs(u_L - u_R) < tol && fabs(p_L - p_R) < tol) {   /* :124-130 */
        *P_star = 0.5 * (p_L + p_R);
        *U_star = 0.5 * (u_L + u_R);
        if (n_iter) *n_iter = -1;
        return fabs(u_L - u_R);
    }

    int n = 0, extra = 0;   /* extra: the last step was the converge rule's extra step */
    for (;;) {                                              /* :133-203 */
        if (!((gap > tol) && (n != N))) {
            /* the converge rule: an iterate within tol that a starting value or an
             * ordinary step gave gets one more step when none was taken or when
             * its gap is above the rounding floor of the velocities or puts p
             * more than 1e-13 off the root */
            if (conv && !extra && n != N &&
                (n == 0 || gap > fmin(0x1p-49 * (fabs(v_L) + fabs(v_R) + c_L + c_R), 1e-13 * p_INT * fabs(k1 - k3))))
                extra = 1;
            else
                break;
        } else {
            extra = 0;
        }
        double t1, t2, t3, dp;
        if (p_INT > p_L) {
            dp = p_INT - p_L;
            t1 = 1.0 / sqrt(1.0 + nuL * dp / p_L);
            t2 = c_L / gL / p_L;
            t3 = 0.5 * t2 * nuL / p_L;
            k1 = t3 * dp * pow(t1, 3.0) - t2 * t1;
        } else {
            t2 = c_L / gL / p_L;
            t1 = 1.0 / pow(p_INT / p_L, nuL);
            k1 = -t1 * t2;
        }
        if (p_INT > p_R) {
            dp = p_INT - p_R;
            t1 = 1.0 / sqrt(1.0 + nuR * dp / p_R);
            t2 = c_R / gR / p_R;
            t3 = 0.5 * t2 * nuR / p_R;
            k3 = t2 * t1 - t3 * dp * pow(t1, 3.0);
        } else {
            t2 = c_R / gR / p_R;
            t1 = 1.0 / pow(p_INT / p_R, nuR);
            k3 = t1 * t2;
        }
        const double p_new = p_INT + (v_R - v_L) / (k1 - k3);
        p_INT = (p_new < 0.0) ? 0.5 * p_INT : p_new;
        ++n;
        v_L = wave_curve(p_INT, u_L, p_L, c_L, gL, muL, nuL, -1.0);
        v_R = wave_curve(p_INT, u_R, p_R, c_R, gR, muR, nuR, +1.0);
        gap = fabs(v_L - v_R);
    }

    *P_star = p_INT;
    *U_star = k1 * (v_R - v_L) / (k1 - k3) + v_L;            /* :205 */
    if (n_iter) *n_iter = n;
    return gap;
}

double orc_riemann_exact(double *U_star, double *P_star, double gammaL, double gammaR,
                         double u_L, double u_R, double p_L, double p_R,
                         double c_L, double c_R, int CRW[2],
                         double eps, double tol, int N, int *n_iter)
{
    const double muL = (gammaL - 1.0) / (2.0 * gammaL), nuL = (gammaL + 1.0) / (2.0 * gammaL);
    const double muR = (gammaR - 1.0) / (2.0 * gammaR), nuR = (gammaR + 1.0) / (2.0 * gammaR);
    return riemann_core(U_star, P_star, gammaL, gammaR, muL, nuL, muR, nuR,
                        u_L, u_R, p_L, p_R, c_L, c_R, CRW, eps, tol, N, n_iter, 0);
}

double orc_riemann_exact_rule(double *U_star, double *P_star, double gammaL, double gammaR,
                              double u_L, double u_R, double p_L, double p_R,
                              double c_L, double c_R, int CRW[2],
                              double eps, double tol, int N, int *n_iter, int rule)
{
    const double muL = (gammaL - 1.0) / (2.0 * gammaL), nuL = (gammaL + 1.0) / (2.0 * gammaL);
    const double muR = (gammaR - 1.0) / (2.0 * gammaR), nuR = (gammaR + 1.0) / (2.0 * gammaR);
    return riemann_core(U_star, P_star, gammaL, gammaR, muL, nuL, muR, nuR,
                        u_L, u_R, p_L, p_R, c_L, c_R, CRW, eps, tol, N, n_iter, rule == 1);
}

/* Newton settings of the interface solves inside the oracle's Q1D (and so its
 * 2D runs): tol <= 0 keeps the reference's tol = eps, rule 0 its stopping rule. */
double orc_newton_tol = 0.0;
int orc_newton_rule = 0;
void orc_set_newton(double tol, int rule)
{
    orc_newton_tol = tol;
    orc_newton_rule = rule;
}

double orc_riemann_exact_ben(double *U_star, double *P_star, double gamma,
                             double u_L, double u_R, double p_L, double p_R,
                             double c_L, double c_R, int CRW[2],
                             double eps, double tol, int N, int *n_iter)
{
    const double mu = (gamma - 1.0) / (2.0 * gamma), nu = (gamma + 1.0) / (2.0 * gamma);
    return riemann_core(U_star, P_star, gamma, gamma, mu, nu, mu, nu,
                        u_L, u_R, p_L, p_R, c_L, c_R, CRW, eps, tol, N, n_iter, 0);
}

/* GuessP (riemann_solver_starPU.c:20-58): the PVRS pressure, kept where the
 * pressure ratio is at most 2 and it lies between the two pressures, or where
 * it lies below both (the two-rarefaction branch is commented out there);
 * otherwise the two-shock estimate; both clipped below at eps. */
double orc_guessp(double DL, double DR, double UL, double UR, double PL, double PR,
                  double CL, double CR, double GammaL, double GammaR, double eps)
{
    const double quser = 2.0;
    const double cup = 0.25 * (DL + DR) * (CL + CR);
    double ppv = 0.5 * (PL + PR) + 0.5 * (UL - UR) * cup;
    ppv = eps > ppv ? eps : ppv;
    const double pmin = PL < PR ? PL : PR;
    const double pmax = PL > PR ? PL : PR;
    const double qmax = pmax / pmin;
    if ((qmax <= quser && pmin <= ppv && ppv <= pmax) || ppv < pmin)
        return ppv;
    const double gel = sqrt((2. / ((GammaL + 1.) * DL)) / ((GammaL - 1.) / (GammaL + 1.) * PL + ppv));
    const double ger = sqrt((2. / ((GammaR + 1.) * DR)) / ((GammaR - 1.) / (GammaR + 1.) * PR + ppv));
    const double pm = (gel * PL + ger * PR - (UR - UL)) / (gel + ger);
    return eps > pm ? eps : pm;
}

/* libm's pow over arrays: the pow the reference's arithmetic uses (glibc), for
 * checking the device pow */
void orc_pow_many(const double *x, const double *y, int n, double *out)
{
    for (int k = 0; k < n; ++k)
        out[k] = pow(x[k], y[k]);
}
