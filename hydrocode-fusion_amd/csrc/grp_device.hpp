// grp_device.hpp -- device-side GRP interface solver for gfx950 (FP64).
//
// One interface = exact Riemann solve for the star state (Newton on the
// velocity gap) + the quasi-1D direct Eulerian GRP time derivatives + the 2D
// flux at t + tau/2 and the interface values at t + tau.  The arithmetic
// follows the reference's expression order exactly so that, with
// -ffp-contract=off, IEEE division/sqrt and an accurate pow, results agree with
// the reference's CPU build to the last bit or the last few ulps:
//
//   riemann_exact  <- Riemann_solver_exact        src/riemann_solver/riemann_solver_exact_Ben.c:30-211
//   grp_q1d        <- linear_GRP_solver_Edir_Q1D  src/riemann_solver/linear_grp_solver_Edir_Q1D.c:38-554
//   grp_2d_flux    <- GRP_2D_flux                 src/flux_calc/flux_solver.c:158-237
//   star_dire_bad  <- star_dire_check (dim 2)     src/inter_process/fluid_var_check.c:65-131
//   minmod2/3      <- src/include/tools.h:44-73
//
// The split path always has lambda = 0 and tangential derivatives +-0; they
// enter as compile-time constants, so hipcc folds every operation on them that
// IEEE arithmetic allows and keeps the rest (x * 0.0 is not x).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "fastpow.hpp"

namespace hgrp {

// Interface side state: value and normal derivative of rho, u, v, p.
struct Side {
    double rho, u, v, p;
    double d_rho, d_u, d_v, d_p;
};

// Constants of one gamma-law gas, formed once (on the host) with exactly the
// expressions the reference writes inline, so every use sees the same double.
// Passed by value in kernel arguments, they live in scalar registers; the ones
// that are a single operation on g or zeta are formed where they are used
// (the same operation, so the same double), which keeps the struct at ten
// doubles: the sweeps run short of scalar registers.
struct Gas {
    double g;                      // gamma
    double mu, nu;                 // (g-1)/(2g), (g+1)/(2g)       wave curves
    double z;                      // zeta = (g-1)/(g+1)
    double inv_g, e_cs;            // 1/g, 0.5 (g-1)/g            star density / sound speed
    double e_sonic, e_a, e_b;      // 2g/(g-1), 0.5/zeta, (1+zeta)/zeta
    double g_gm1;                  // g/(g-1)                      energy flux
    __host__ __device__ double gm1() const { return g - 1.0; }
    __host__ __device__ double gp1() const { return g + 1.0; }
    __host__ __device__ double opz() const { return 1.0 + z; }          // 1 + zeta
    __host__ __device__ double op2z() const { return 1.0 + 2.0 * z; }   // 1 + 2 zeta
    __host__ __device__ double omz() const { return 1 - z; }            // 1 - zeta
    __host__ __device__ double tpz() const { return 2.0 + z; }          // 2 + zeta
    __host__ __device__ double zz() const { return z * z; }             // zeta^2
    __host__ __device__ double omzz() const { return 1.0 - z * z; }     // 1 - zeta^2
    __host__ __device__ double zzm1() const { return z * z - 1.0; }     // zeta^2 - 1
};

__host__ __device__ inline Gas make_gas(double g)
{
    Gas s;
    s.g = g;
    s.mu = (g - 1.0) / (2.0 * g);
    s.nu = (g + 1.0) / (2.0 * g);
    s.z = (g - 1.0) / (g + 1.0);
    s.inv_g = 1.0 / g;
    s.e_cs = 0.5 * (g - 1.0) / g;
    s.e_sonic = 2.0 * g / (g - 1.0);
    s.e_a = 0.5 / s.z;
    s.e_b = (1.0 + s.z) / s.z;
    s.g_gm1 = g / (g - 1.0);
    return s;
}

// a / b for a divisor b known when the kernel starts, from y = RN(1/b)
// (computed once): q = a y, r = a - q b (exact in an FMA), q + r y rounded.
// For |b| in [2^-100, 2^100] and |q| in [2^-800, 2^800] this is the correctly
// rounded quotient (Markstein's theorem: y correctly rounded, q within one ulp,
// no underflow in r), i.e. bit for bit what IEEE a / b gives; a = +-0 gives
// q = a y = +-0 exactly.  Anything else (zeros of b, infinities, NaN, the
// extremes of the range) takes the division itself.
struct RecipDiv {
    double b, y;
    bool ok;
};
__host__ __device__ inline RecipDiv recip_div(double b)
{
    RecipDiv d;
    d.b = b;
    d.y = 1.0 / b;
    const double ab = fabs(b);
    d.ok = ab >= 0x1p-100 && ab <= 0x1p100;
    return d;
}
// the division itself, out of line: div_by is inlined at dozens of sites and
// this path (zeros, infinities, NaN, extreme quotients) is rarely taken
__device__ __attribute__((noinline)) static double div_slow(double a, double b) { return a / b; }

__device__ __forceinline__ double div_by(double a, const RecipDiv &d)
{
    const double q = a * d.y;
    const double r = fma(-q, d.b, a);
    const double q2 = fma(r, d.y, q);
    const double aq = fabs(q);
    if (__builtin_expect(d.ok && aq >= 0x1p-800 && aq <= 0x1p800, 1)) return q2;
    if (a == 0.0 && d.ok) return q;
    return div_slow(a, d.b);
}

__device__ __forceinline__ double minmod2(double s_L, double s_R)
{
    if (s_L * s_R <= 0.0) return 0.0;
    if (s_R > 0.0 && s_R < s_L) return s_R;
    if (s_R <= 0.0 && s_R > s_L) return s_R;
    return s_L;
}

__device__ __forceinline__ double minmod3(double s_L, double s_R, double s_m)
{
    if (s_L * s_m <= 0.0 || s_R * s_m <= 0.0) return 0.0;
    if (s_m > 0.0 && s_m < s_L && s_m < s_R) return s_m;
    if (s_m <= 0.0 && s_m > s_L && s_m > s_R) return s_m;
    if (s_R > 0.0 && s_R < s_L) return s_R;
    if (s_R <= 0.0 && s_R > s_L) return s_R;
    return s_L;
}

// x^y, correctly rounded in all but ~1e-5 of cases (fastpow.hpp); powers of
// one base share its logarithm through PowBase.
__device__ __forceinline__ double gpow(double x, double y) { return fast_pow(x, y); }

// velocity on the 1-wave (sign -1) or 3-wave (sign +1) curve through (u_s, p_s)
// at pressure p; on the rarefaction branch *base receives log(p / p_s) for the
// Newton slope at the same p
template <int SIGN>
__device__ __forceinline__ double wave_curve(double p, double u_s, double p_s, double c_s, const Gas &G,
                                             DD *logx)
{
    double w;
    if (p > p_s) {
        const double dp = p - p_s;
        const double s = sqrt(1.0 + G.nu * dp / p_s);
        w = dp * c_s / G.g / p_s / s;
    } else {
        const PowBase b = pow_base(p / p_s);
        *logx = b.L;
        w = pow_with(b, G.mu) - 1.0;
        w = 2.0 * c_s * w / G.gm1();
    }
    return SIGN < 0 ? u_s - w : u_s + w;
}

// x^y for x = p/p_s on the rarefaction branch, given log(x) from wave_curve
__device__ __forceinline__ double pow_from_log(double x, DD L, double y)
{
    PowBase b;
    b.x = x;
    b.ok = pow_base_ok(x);
    b.L = L;
    return pow_with(b, y);
}

// Newton steps allowed from a learned starting guess before the solve restarts
// from the reference's guess.
constexpr int GUESS_ITERS = 3;

// The converge rule (not in the reference; its loop stops at the first iterate
// whose velocity gap is within tol).  An absolute gap of tol leaves p up to
// tol / (p |dv/dp|) relative off the root, which near vacuum (p |dv/dp| down
// to ~1e-3) is 1e-10 at tol = 1e-12, and even on weak waves leaves ~1e-13:
// where the stop falls then depends on the starting guess.  Under this rule
// an iterate within tol gets one more Newton step if none has been taken yet,
// or if its gap is above the rounding floor of the curves' velocities,
// CONV_ULP (|v_L| + |v_R| + c_L + c_R) (8 ulp), or above CONV_REL p |k1 - k3|
// (the last tangents put p more than CONV_REL off the root; the smaller
// threshold decides, so that near vacuum, where p |dv/dp| is small, p itself is
// converged); an iterate reached by that extra step ends the loop if it meets
// tol (so the count stays bounded where p* is ill-conditioned), else the
// ordinary steps go on.  Quadratic convergence takes an error of e to ~e^2, so
// either starting guess ends at the root to within rounding.
constexpr double CONV_ULP = 0x1p-49, CONV_REL = 1e-13;

struct StarResult {
    double u_star, p_star;
    double gap;    // the solver's return value: |v_L - v_R| at the end (|u_L - u_R| on the early exit)
    int crw0, crw1;
    int iters;     // Newton iterations taken (-1: equal-state early exit)
};

// The update of one Newton step of Riemann_solver_exact
// (riemann_solver_exact_Ben.c:133-203): tangents k1, k3 at p, the new p
// (halved when negative).  The wave curves at the new p follow in the caller.
__device__ __forceinline__ void newton_update(const Gas &GL, const Gas &GR, double p_L, double p_R, double c_L,
                                              double c_R, double &p_INT, double v_L, double v_R, double &k1,
                                              double &k3, const DD &bL, const DD &bR)
{
    const double gL = GL.g, gR = GR.g;
    double t1, t2, t3, dp;
    if (p_INT > p_L) {
        dp = p_INT - p_L;
        t1 = 1.0 / sqrt(1.0 + GL.nu * dp / p_L);
        t2 = c_L / gL / p_L;
        t3 = 0.5 * t2 * GL.nu / p_L;
        k1 = t3 * dp * cube(t1) - t2 * t1;
    } else {
        t2 = c_L / gL / p_L;
        t1 = 1.0 / pow_from_log(p_INT / p_L, bL, GL.nu);   // bL = log(p_INT / p_L) from the curve value
        k1 = -t1 * t2;
    }
    if (p_INT > p_R) {
        dp = p_INT - p_R;
        t1 = 1.0 / sqrt(1.0 + GR.nu * dp / p_R);
        t2 = c_R / gR / p_R;
        t3 = 0.5 * t2 * GR.nu / p_R;
        k3 = t2 * t1 - t3 * dp * cube(t1);
    } else {
        t2 = c_R / gR / p_R;
        t1 = 1.0 / pow_from_log(p_INT / p_R, bR, GR.nu);
        k3 = t1 * t2;
    }
    const double p_new = p_INT + (v_R - v_L) / (k1 - k3);
    p_INT = (p_new < 0.0) ? 0.5 * p_INT : p_new;
}

// GuessP (src/riemann_solver/riemann_solver_starPU.c:20-58), the reference's
// PVRS / two-shock starting pressure: PVRS where the pressure ratio is at most
// 2 and the estimate lies between the two pressures, or lies below both;
// otherwise the two-shock estimate; clipped below at eps.  The 2D path never
// calls it (riemann_exact starts from the halved tangent intersection); it is
// the named alternative starting guess (hgrp_riemann_pvrs_batch).
__device__ __forceinline__ double guess_pvrs(double DL, double DR, double UL, double UR, double PL, double PR,
                                             double CL, double CR, double GammaL, double GammaR, double eps)
{
    const double cup = 0.25 * (DL + DR) * (CL + CR);
    double ppv = 0.5 * (PL + PR) + 0.5 * (UL - UR) * cup;
    ppv = eps > ppv ? eps : ppv;
    const double pmin = PL < PR ? PL : PR, pmax = PL > PR ? PL : PR;
    const double qmax = pmax / pmin;
    if ((qmax <= 2.0 && pmin <= ppv && ppv <= pmax) || ppv < pmin) return ppv;
    const double gel = sqrt((2. / ((GammaL + 1.) * DL)) / ((GammaL - 1.) / (GammaL + 1.) * PL + ppv));
    const double ger = sqrt((2. / ((GammaR + 1.) * DR)) / ((GammaR - 1.) / (GammaR + 1.) * PR + ppv));
    const double pm = (gel * PL + ger * PR - (UR - UL)) / (gel + ger);
    return eps > pm ? eps : pm;
}

// Riemann_solver_exact: wave types, the halved acoustic-tangent guess, Newton
// on the velocity gap, u* from the last tangents.
// p_guess > 0 (not in the reference): first up to GUESS_ITERS Newton steps
// from p_guess; if they meet tol that is the answer, otherwise the solve
// restarts from the reference's guess exactly as the reference does.  iters
// counts both phases.  CONV: the converge rule (CONV_REL) in both phases.
template <bool CONV = false>
__device__ __forceinline__ StarResult riemann_exact(const Gas &GL, const Gas &GR, double u_L, double u_R,
                                                    double p_L, double p_R, double c_L, double c_R,
                                                    double eps, double tol, int N, double p_guess = -1.0)
{
    const double gL = GL.g, gR = GR.g;
    StarResult r;
    DD bL = {0.0, 0.0}, bR = {0.0, 0.0};
    const double u_LR = wave_curve<-1>(p_R, u_L, p_L, c_L, GL, &bL);
    const double u_RL = wave_curve<+1>(p_L, u_R, p_R, c_R, GR, &bR);
    r.crw1 = (u_LR > u_R + eps) ? 0 : 1;
    r.crw0 = (u_RL > u_L - eps) ? 1 : 0;

    const double k1_0 = -c_L / p_L / gL;
    const double k3_0 = c_R / p_R / gR;
    double k1 = k1_0, k3 = k3_0;
    double p_INT = (k1 * p_L - k3 * p_R - u_L + u_R) / (k1 - k3);
    if (p_INT < 0)
        p_INT = (p_L < p_R) ? p_L : p_R;
    p_INT = 0.5 * p_INT;

    if (fabs(u_L - u_R) < tol && fabs(p_L - p_R) < tol) {
        r.p_star = 0.5 * (p_L + p_R);
        r.u_star = 0.5 * (u_L + u_R);
        r.gap = fabs(u_L - u_R);
        r.iters = -1;
        return r;
    }
    // One loop for both phases, so the Newton code is laid out once: st 0
    // evaluates the curves at the learned guess, 1 steps from it, 2 evaluates
    // them at the reference's guess p_INT (the restart, tangents reset), 3 steps
    // from it exactly as the reference's loop does; 4 and 5 are the converge
    // rule's extra step in the learned phase and in the reference's phase.
    const bool guessed = p_guess > 0.0;
    int st = guessed ? 0 : 2, n = 0, done = 0;
    double p = 0.0, v_L = 0.0, v_R = 0.0, gap = 0.0;
    for (;;) {
        if (st == 0 || st == 2) {
            p = st == 0 ? p_guess : p_INT;
            k1 = k1_0;
            k3 = k3_0;
        } else {
            newton_update(GL, GR, p_L, p_R, c_L, c_R, p, v_L, v_R, k1, k3, bL, bR);
        }
        v_L = wave_curve<-1>(p, u_L, p_L, c_L, GL, &bL);
        v_R = wave_curve<+1>(p, u_R, p_R, c_R, GR, &bR);
        gap = fabs(v_L - v_R);
        const bool learned_phase = st <= 1 || st == 4;
        if (st == 1 || st == 4) ++done;
        if (st == 3 || st == 5) ++n;
        // converge rule: an iterate within tol, reached from a starting value or by an ordinary
        // step, gets one more step if none was taken or if the error is above CONV_REL
        const bool extra = CONV && st != 4 && st != 5 &&
                           (st == 0 || st == 2 ||
                            gap > fmin(CONV_ULP * (fabs(v_L) + fabs(v_R) + c_L + c_R), CONV_REL * p * fabs(k1 - k3)));
        if (learned_phase) {
            if (gap <= tol) {
                if (extra) {
                    st = 4;
                    continue;
                }
                r.p_star = p;
                r.u_star = k1 * (v_R - v_L) / (k1 - k3) + v_L;
                r.gap = gap;
                r.iters = done;
                return r;
            }
            // a NaN gap restarts from the reference's guess, as an unmet tolerance does
            st = (gap > tol && done < GUESS_ITERS) ? 1 : 2;
        } else {
            // the reference's loop condition: a NaN gap ends the loop
            if (gap > tol && n != N) {
                st = 3;
            } else if (extra && n != N) {
                st = 5;
            } else {
                break;
            }
        }
    }
    r.p_star = p;
    r.u_star = k1 * (v_R - v_L) / (k1 - k3) + v_L;
    r.gap = gap;
    r.iters = n + done;
    return r;
}

// Upwind derivative when the t-axis lies outside the wave fan.
__device__ __forceinline__ void side_trivial(double D[4], double U[4], double rho, double u, double v,
                                             double p, double d_rho, double d_u, double d_v, double d_p,
                                             double t_rho, double t_u, double t_v, double t_p, double c,
                                             double lu, double lv)
{
    U[0] = rho; U[1] = u; U[2] = v; U[3] = p;
    D[0] = -(u - lu) * d_rho - (v - lv) * t_rho - rho * (d_u + t_v);
    D[1] = -(u - lu) * d_u   - (v - lv) * t_u   - d_p / rho;
    D[2] = -(u - lu) * d_v   - (v - lv) * t_v   - t_p / rho;
    D[3] = -(u - lu) * d_p   - (v - lv) * t_p   - rho * c * c * (d_u + t_v);
}

// linear_GRP_solver_Edir_Q1D, single-fluid members, with atc = eps.
// TL/TR: tangential derivatives (rho, u, v, p) of the two sides.
// tol: Newton tolerance (the reference passes eps); p_guess: see riemann_exact.
// TF (two fluids, the reference's MULTIFLUID_BASICS members of Q1D, :55-61 and
// the U[4]/D[4] lines): FA = {a_L, d_a_L, a_R, d_a_R}, the volume fraction of
// fluid a and its normal derivative on each side; FO receives its value on the
// t-axis and its time derivative (carried with the contact like v; the split
// path's tangential derivatives of a are zero).
// The rest of linear_GRP_solver_Edir_Q1D's interface, for callers that use it
// as a function on its own (the solver's C-ABI entry points): atc and the
// grid velocity (lambda_u, lambda_v), which the flux path fixes at eps and 0;
// wave_speed[2] and U_star[6] out.  The sweeps pass none.  CONV: the
// converge rule of the exact solve (riemann_exact).
struct Q1DExtra {
    double atc, lu, lv;
    double *ws, *us;
};

template <bool TF = false, bool CONV = false>
__device__ __forceinline__ void grp_q1d(double D[4], double U[4], const Side &L, const Side &R,
                                        const double TL[4], const double TR[4],
                                        const Gas &GL, const Gas &GR, double eps, double tol,
                                        double p_guess, const double *FA = nullptr, double *FO = nullptr,
                                        const Q1DExtra *X = nullptr)
{
    const double gL = GL.g, gR = GR.g;
    const double lu = X ? X->lu : 0.0, lv = X ? X->lv : 0.0;
    const double atc = X ? X->atc : eps;
    const double rho_L = L.rho, rho_R = R.rho, u_L = L.u, u_R = R.u;
    const double v_L = L.v, v_R = R.v, p_L = L.p, p_R = R.p;
    const double d_rho_L = L.d_rho, d_rho_R = R.d_rho, d_u_L = L.d_u, d_u_R = R.d_u;
    const double d_v_L = L.d_v, d_v_R = R.d_v, d_p_L = L.d_p, d_p_R = R.d_p;
    const double t_rho_L = TL[0], t_u_L = TL[1], t_v_L = TL[2], t_p_L = TL[3];
    const double t_rho_R = TR[0], t_u_R = TR[1], t_v_R = TR[2], t_p_R = TR[3];
    const double zL = GL.z, zR = GR.z;

    int crw0 = 0, crw1 = 0;
    double u_star, p_star, rs_L, rs_R, cs_L, cs_R, speed_L, speed_R, C;

    const double c_L = sqrt(gL * p_L / rho_L);
    const double c_R = sqrt(gR * p_R / rho_R);
    const double dist = sqrt((rho_L - rho_R) * (rho_L - rho_R) + (u_L - u_R) * (u_L - u_R)
                             + (p_L - p_R) * (p_L - p_R));

    if (dist < atc && atc < 2 * eps) {
        u_star = 0.5 * (u_R + u_L);
        p_star = 0.5 * (p_R + p_L);
        rs_L = rho_L; cs_L = c_L; speed_L = u_star - cs_L;
        rs_R = rho_R; cs_R = c_R; speed_R = u_star + cs_R;
    } else {
        const StarResult st = riemann_exact<CONV>(GL, GR, u_L, u_R, p_L, p_R, c_L, c_R, eps, tol, 500, p_guess);
        u_star = st.u_star; p_star = st.p_star; crw0 = st.crw0; crw1 = st.crw1;
        if (crw0) {
            const PowBase b = pow_base(p_star / p_L);
            rs_L = rho_L * pow_with(b, GL.inv_g);
            cs_L = c_L * pow_with(b, GL.e_cs);
            speed_L = u_L - c_L;
        } else {
            rs_L = rho_L * (p_star + zL * p_L) / (p_L + zL * p_star);
            cs_L = sqrt(gL * p_star / rs_L);
            speed_L = u_L - c_L * sqrt(0.5 * (GL.gp1() * (p_star / p_L) + GL.gm1()) / gL);
        }
        if (crw1) {
            const PowBase b = pow_base(p_star / p_R);
            rs_R = rho_R * pow_with(b, GR.inv_g);
            cs_R = c_R * pow_with(b, GR.e_cs);
            speed_R = u_R + c_R;
        } else {
            rs_R = rho_R * (p_star + zR * p_R) / (p_R + zR * p_star);
            cs_R = sqrt(gR * p_star / rs_R);
            speed_R = u_R + c_R * sqrt(0.5 * (GR.gp1() * (p_star / p_R) + GR.gm1()) / gR);
        }
    }

    if (X) {
        X->ws[0] = speed_L; X->ws[1] = speed_R;
        X->us[0] = rs_L; X->us[1] = u_star; X->us[2] = rs_R; X->us[3] = p_star; X->us[4] = cs_L; X->us[5] = cs_R;
    }
    if (speed_L > lu) {
        side_trivial(D, U, rho_L, u_L, v_L, p_L, d_rho_L, d_u_L, d_v_L, d_p_L, t_rho_L, t_u_L, t_v_L, t_p_L, c_L, lu, lv);
        if constexpr (TF) { FO[0] = FA[0]; FO[1] = -(u_L - lu) * FA[1]; }
        return;
    }
    if (speed_R < lu) {
        side_trivial(D, U, rho_R, u_R, v_R, p_R, d_rho_R, d_u_R, d_v_R, d_p_R, t_rho_R, t_u_R, t_v_R, t_p_R, c_R, lu, lv);
        if constexpr (TF) { FO[0] = FA[2]; FO[1] = -(u_R - lu) * FA[3]; }
        return;
    }

    if (dist < atc) {
        // acoustic case.  With atc < 2 eps the star state is the averaged one,
        // which never lies inside a fan (u_star -+ c_star > 0 is exactly the
        // trivial test above); the flux path (X == nullptr) is that case.  A
        // larger atc (INFINITY: acoustic only) takes the exact star state,
        // whose rarefaction fans the t-axis can lie in (:173-192).
        if (X && crw0 && ((u_star - cs_L) > lu)) {
            U[1] = zL * (u_L + 2.0 * (c_L + lu) / GL.gm1());
            C = U[1] - lu;
            U[3] = pow_with(pow_base(C / c_L), GL.e_sonic) * p_L;
            U[0] = gL * U[3] / C / C;
            U[2] = v_L;
        } else if (X && crw1 && ((u_star + cs_R) < lu)) {
            U[1] = zR * (u_R - 2.0 * (c_R - lu) / GR.gm1());
            C = lu - U[1];
            U[3] = pow_with(pow_base(C / c_R), GR.e_sonic) * p_R;
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
        if constexpr (TF) {
            const bool left = u_star > lu;
            FO[0] = left ? FA[0] : FA[2];
            FO[1] = -(U[1] - lu) * (left ? FA[1] : FA[3]);
        }
        return;
    }

    if (crw0 && ((u_star - cs_L) > lu)) {                 // t-axis inside the 1-CRW
        U[1] = zL * (u_L + 2.0 * (c_L + lu) / GL.gm1());
        C = U[1] - lu;
        const PowBase b = pow_base(C / c_L);             // c_frac = C / c_L below
        U[3] = pow_with(b, GL.e_sonic) * p_L;
        U[0] = gL * U[3] / C / C;
        U[2] = v_L;
        const double TdS = (d_p_L - d_rho_L * c_L * c_L) / GL.gm1() / rho_L;
        const double d_Psi = d_u_L + (gL * d_p_L / c_L - c_L * d_rho_L) / GL.gm1() / rho_L;
        const double pw_a = pow_with(b, GL.e_a), pw_b = pow_with(b, GL.e_b);
        D[1] = (GL.opz() * pw_a + zL * pw_b);
        D[1] = D[1] / GL.op2z() * TdS;
        D[1] = D[1] - c_L * pw_a * d_Psi;
        D[3] = U[0] * (U[1] - lu) * D[1];
        D[0] = U[0] * (U[1] - lu) * pw_b * TdS * GL.gm1();
        D[0] = (D[0] + D[3]) / C / C;
        D[2] = -(U[1] - lu) * d_v_L * U[0] / rho_L;
        if constexpr (TF) { FO[0] = FA[0]; FO[1] = -(U[1] - lu) * FA[1] * U[0] / rho_L; }
    } else if (crw1 && ((u_star + cs_R) < lu)) {          // t-axis inside the 3-CRW
        U[1] = zR * (u_R - 2.0 * (c_R - lu) / GR.gm1());
        C = lu - U[1];
        const PowBase b = pow_base(C / c_R);
        U[3] = pow_with(b, GR.e_sonic) * p_R;
        U[0] = gR * U[3] / C / C;
        U[2] = v_R;
        const double TdS = (d_p_R - d_rho_R * c_R * c_R) / GR.gm1() / rho_R;
        const double d_Phi = d_u_R - (gR * d_p_R / c_R - c_R * d_rho_R) / GR.gm1() / rho_R;
        const double pw_a = pow_with(b, GR.e_a), pw_b = pow_with(b, GR.e_b);
        D[1] = (GR.opz() * pw_a + zR * pw_b);
        D[1] = D[1] / GR.op2z() * TdS;
        D[1] = D[1] + c_R * pw_a * d_Phi;
        D[3] = U[0] * (U[1] - lu) * D[1];
        D[0] = U[0] * (U[1] - lu) * pw_b * TdS * GR.gm1();
        D[0] = (D[0] + D[3]) / C / C;
        D[2] = -(U[1] - lu) * d_v_R * U[0] / rho_R;
        if constexpr (TF) { FO[0] = FA[2]; FO[1] = -(U[1] - lu) * FA[3] * U[0] / rho_R; }
    } else {                                              // non-sonic
        double a_L, b_L, d_L, a_R, b_R, d_R;
        if (u_star < lu) {
            U[0] = rs_R; U[1] = u_star; U[2] = v_R; U[3] = p_star; C = cs_R;
        } else {
            U[0] = rs_L; U[1] = u_star; U[2] = v_L; U[3] = p_star; C = cs_L;
        }
        // pow(cs/c, (1+z)/z) is needed again for D[0] below on the CRW side
        double pwL_b = 0.0, pwR_b = 0.0;
        if (crw0) {
            a_L = 1.0;
            b_L = 1.0 / rs_L / cs_L;
            const PowBase b = pow_base(cs_L / c_L);
            const double TdS = (d_p_L - d_rho_L * c_L * c_L) / GL.gm1() / rho_L;
            const double d_Psi = d_u_L + (gL * d_p_L / c_L - c_L * d_rho_L) / GL.gm1() / rho_L;
            const double pw_a = pow_with(b, GL.e_a);
            pwL_b = pow_with(b, GL.e_b);
            d_L = (GL.opz() * pw_a + zL * pwL_b);
            d_L = d_L / GL.op2z() * TdS;
            d_L = d_L - c_L * pw_a * d_Psi;
        } else {
            const double SmUs = -sqrt(0.5 * (GL.gp1() * p_L + GL.gm1() * p_star) / rs_L);
            const double SmUL = -sqrt(0.5 * (GL.gp1() * p_star + GL.gm1() * p_L) / rho_L);
            const double VAR = sqrt(GL.omz() / (rho_L * (p_star + zL * p_L)));
            const double H1 = 0.5 * VAR * (p_star + GL.op2z() * p_L) / (p_star + zL * p_L);
            const double H2 = -0.5 * VAR * (GL.tpz() * p_star + zL * p_L) / (p_star + zL * p_L);
            const double H3 = -0.5 * VAR * (p_star - p_L) / rho_L;
            const double L_p = -1.0 / rho_L - SmUL * H2;
            const double L_u = SmUL + rho_L * (c_L * c_L * H2 + H3);
            const double L_rho = -SmUL * H3;
            a_L = 1.0 - rs_L * SmUs * H1;
            b_L = -SmUs / (rs_L * cs_L * cs_L) + H1;
            d_L = L_rho * d_rho_L + L_u * d_u_L + L_p * d_p_L;
        }
        if (crw1) {
            a_R = 1.0;
            b_R = -1.0 / rs_R / cs_R;
            const PowBase b = pow_base(cs_R / c_R);
            const double TdS = (d_p_R - d_rho_R * c_R * c_R) / GR.gm1() / rho_R;
            const double d_Phi = d_u_R - (gR * d_p_R / c_R - c_R * d_rho_R) / GR.gm1() / rho_R;
            const double pw_a = pow_with(b, GR.e_a);
            pwR_b = pow_with(b, GR.e_b);
            d_R = (GR.opz() * pw_a + zR * pwR_b);
            d_R = d_R / GR.op2z() * TdS;
            d_R = d_R + c_R * pw_a * d_Phi;
        } else {
            const double SmUs = sqrt(0.5 * (GR.gp1() * p_R + GR.gm1() * p_star) / rs_R);
            const double SmUR = sqrt(0.5 * (GR.gp1() * p_star + GR.gm1() * p_R) / rho_R);
            const double VAR = sqrt(GR.omz() / (rho_R * (p_star + zR * p_R)));
            const double H1 = 0.5 * VAR * (p_star + GR.op2z() * p_R) / (p_star + zR * p_R);
            const double H2 = -0.5 * VAR * (GR.tpz() * p_star + zR * p_R) / (p_star + zR * p_R);
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
            if (crw1) {
                D[0] = rs_R * (u_star - lu) * pwR_b * (d_p_R - d_rho_R * c_R * c_R) / rho_R;
                D[0] = (D[0] + D[3]) / cs_R / cs_R;
                D[2] = -U[1] * d_v_R * U[0] / rho_R;
                D[2] = D[2] + lu * d_v_R;
                if constexpr (TF) { FO[0] = FA[2]; FO[1] = -U[1] * FA[3] * U[0] / rho_R + lu * FA[3]; }
            } else {
                const double SmUs = sqrt(0.5 * (GR.gp1() * p_R + GR.gm1() * p_star) / rs_R);
                const double SmUR = sqrt(0.5 * (GR.gp1() * p_star + GR.gm1() * p_R) / rho_R);
                const double VAR = p_R + zR * p_star;
                const double H1 = rho_R * p_R * GR.omzz() / VAR / VAR;
                const double H2 = rho_R * p_star * GR.zzm1() / VAR / VAR;
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
                if constexpr (TF) { FO[0] = FA[2]; FO[1] = -U[1] * SmUR * FA[3] / SmUs + lu * FA[3]; }
            }
        } else {
            if (crw0) {
                D[0] = rs_L * (u_star - lu) * pwL_b * (d_p_L - d_rho_L * c_L * c_L) / rho_L;
                D[0] = (D[0] + D[3]) / cs_L / cs_L;
                D[2] = -U[1] * d_v_L * U[0] / rho_L;
                D[2] = D[2] + lu * d_v_L;
                if constexpr (TF) { FO[0] = FA[0]; FO[1] = -U[1] * FA[1] * U[0] / rho_L + lu * FA[1]; }
            } else {
                const double SmUs = -sqrt(0.5 * (GL.gp1() * p_L + GL.gm1() * p_star) / rs_L);
                const double SmUL = -sqrt(0.5 * (GL.gp1() * p_star + GL.gm1() * p_L) / rho_L);
                const double VAR = p_L + zL * p_star;
                const double H1 = rho_L * p_L * GL.omzz() / VAR / VAR;
                const double H2 = rho_L * p_star * GL.zzm1() / VAR / VAR;
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
                if constexpr (TF) { FO[0] = FA[0]; FO[1] = -U[1] * SmUL * FA[1] / SmUs + lu * FA[1]; }
            }
        }
    }
    // transversal terms
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

// star_dire_check (dim 2): 0 ok, 1 non-positive, 2 non-finite state, 3 non-finite derivative
__device__ __forceinline__ int star_dire_bad(const double mid[4], const double dire[4], double eps)
{
    if (mid[3] < eps || mid[0] < eps) return 1;
    if (!isfinite(mid[1]) || !isfinite(mid[2]) || !isfinite(mid[0]) || !isfinite(mid[3])) return 2;
    if (!isfinite(dire[1]) || !isfinite(dire[2]) || !isfinite(dire[0]) || !isfinite(dire[3])) return 3;
    return 0;
}

// Fluxes over tau and interface values at t + tau from the Q1D mid state and
// its time derivative, rotated back to grid axes (flux_solver.c:190-236);
// star_dire_check failures return their code with zero F and I.  The normal
// (n_x, n_y) is (1, 0) or (0, 1); a compile-time constant in the split sweeps.
__device__ __forceinline__ int flux_from_mid_n(const double mid[4], const double dire[4], const Gas &G, double n_x,
                                               double n_y, double tau, double eps, double F[4], double I[4])
{
    const int bad = star_dire_bad(mid, dire, eps);
    if (bad) {
        F[0] = F[1] = F[2] = F[3] = 0.0;
        I[0] = I[1] = I[2] = I[3] = 0.0;
        return bad;
    }
    const double rho_mid = mid[0] + 0.5 * tau * dire[0];
    const double u_mid = (mid[1] + 0.5 * tau * dire[1]) * n_x - (mid[2] + 0.5 * tau * dire[2]) * n_y;
    const double v_mid = (mid[1] + 0.5 * tau * dire[1]) * n_y + (mid[2] + 0.5 * tau * dire[2]) * n_x;
    const double p_mid = mid[3] + 0.5 * tau * dire[3];
    F[0] = rho_mid * (u_mid * n_x + v_mid * n_y);
    F[1] = F[0] * u_mid + p_mid * n_x;
    F[2] = F[0] * v_mid + p_mid * n_y;
    F[3] = G.g_gm1 * p_mid / rho_mid + 0.5 * (u_mid * u_mid + v_mid * v_mid);
    F[3] = F[0] * F[3];
    I[1] = (mid[1] + tau * dire[1]) * n_x - (mid[2] + tau * dire[2]) * n_y;
    I[2] = (mid[1] + tau * dire[1]) * n_y + (mid[2] + tau * dire[2]) * n_x;
    I[0] = mid[0] + tau * dire[0];
    I[3] = mid[3] + tau * dire[3];
    return 0;
}
template <int DIR>
__device__ __forceinline__ int flux_from_mid(const double mid[4], const double dire[4], const Gas &G, double tau,
                                             double eps, double F[4], double I[4])
{
    return flux_from_mid_n(mid, dire, G, DIR == 0 ? 1.0 : 0.0, DIR == 0 ? 0.0 : 1.0, tau, eps, F, I);
}

// GRP_2D_flux for the axis normal DIR (0: x, n = (1,0); 1: y, n = (0,1)) with
// the split path's zero tangential derivatives (+0.0 in x, -0.0 in y).
// On success F = (F_rho, F_u, F_v, F_e), I = (rho, u, v, p) at t + tau; on a
// star/derivative failure F and I are zero and the check code is returned.
// tol, p_guess, CONV: Newton tolerance, learned starting pressure and the
// converge rule (riemann_exact); the reference's path is tol = eps,
// p_guess <= 0, CONV false.
template <int DIR, bool CONV = false>
__device__ __forceinline__ int grp_2d_flux(Side L, Side R, const Gas &G, double tau, double eps,
                                           double F[4], double I[4], double tol, double p_guess = -1.0)
{
    constexpr double n_x = DIR == 0 ? 1.0 : 0.0;
    constexpr double n_y = DIR == 0 ? 0.0 : 1.0;
    constexpr double tz = DIR == 0 ? 0.0 : -0.0;
    const double u = L.u * n_x + L.v * n_y;
    const double u_R = R.u * n_x + R.v * n_y;
    const double d_u = L.d_u * n_x + L.d_v * n_y;
    const double d_u_R = R.d_u * n_x + R.d_v * n_y;
    constexpr double t_u = tz * n_x + tz * n_y;
    constexpr double t_v = -tz * n_y + tz * n_x;
    L.v = -L.u * n_y + L.v * n_x;
    R.v = -R.u * n_y + R.v * n_x;
    L.d_v = -L.d_u * n_y + L.d_v * n_x;
    R.d_v = -R.d_u * n_y + R.d_v * n_x;
    L.u = u; R.u = u_R; L.d_u = d_u; R.d_u = d_u_R;
    const double T[4] = {tz, t_u, t_v, tz};

    double dire[4], mid[4];
    grp_q1d<false, CONV>(dire, mid, L, R, T, T, G, G, eps, tol, p_guess);
    return flux_from_mid<DIR>(mid, dire, G, tau, eps, F, I);
}

// GRP_2D_flux with the transversal derivatives of both sides and a gas per
// side (GRP_2D_flux itself, flux_solver.c:158-237: any unit normal (n_x, n_y),
// the energy flux with the left side's gamma).  TL, TR hold (t_rho, t_u, t_v,
// t_p) in grid axes; they are rotated with the velocities (:167-185).
template <bool CONV = false>
__device__ __forceinline__ int grp_2d_flux_g(Side L, Side R, const double TL[4], const double TR[4], const Gas &GL,
                                             const Gas &GR, double n_x, double n_y, double tau, double eps,
                                             double F[4], double I[4], double tol, double p_guess = -1.0)
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
    double dire[4], mid[4];
    grp_q1d<false, CONV>(dire, mid, L, R, TLr, TRr, GL, GR, eps, tol, p_guess);
    return flux_from_mid_n(mid, dire, GL, n_x, n_y, tau, eps, F, I);
}

// The unsplit step's form: one gas, normal (1, 0) or (0, 1)
// (flux_generator_{x,y}.c:96-121 with Transversal = true).
template <bool CONV = false>
__device__ __forceinline__ int grp_2d_flux_n(Side L, Side R, const double TL[4], const double TR[4], const Gas &G,
                                             double n_x, double n_y, double tau, double eps, double F[4], double I[4],
                                             double tol, double p_guess = -1.0)
{
    return grp_2d_flux_g<CONV>(L, R, TL, TR, G, G, n_x, n_y, tau, eps, F, I, tol, p_guess);
}

// Two gamma-law fluids a and b mixed by the volume fraction a of fluid a:
// gamma = 1/(a/(gamma_a-1) + (1-a)/(gamma_b-1)) + 1
// (src/inter_process_unstruct/cons_qty_calc.c:70).  Pure cells use their
// fluid's constants as they are, so a = 1 everywhere is the single-fluid path.
struct TwoGas {
    Gas a, b;
};
__device__ __forceinline__ double mix_gamma(double fa, const TwoGas &T)
{
    return 1.0 / (fa / (T.a.g - 1.0) + (1.0 - fa) / (T.b.g - 1.0)) + 1.0;
}
__device__ __forceinline__ Gas gas_of(double fa, const TwoGas &T)
{
    if (fa == 1.0) return T.a;
    if (fa == 0.0) return T.b;
    return make_gas(mix_gamma(fa, T));
}

// GRP_2D_flux with the MULTIFLUID_BASICS members (src/flux_calc/flux_solver.c:158-237,
// :203-235): gamma per side from the reconstructed volume fractions, gamma at
// t + tau/2 from a_mid for the energy flux.  FA = {a_L, d_a_L, a_R, d_a_R};
// on success FZ = {u_n a at t + tau/2, u_n at t + tau/2, a at t + tau}.
template <int DIR, bool CONV = false>
__device__ __forceinline__ int grp_2d_flux_tf(Side L, Side R, const TwoGas &T, const double FA[4], double tau,
                                              double eps, double F[4], double I[4], double FZ[3], double tol,
                                              double p_guess = -1.0)
{
    constexpr double n_x = DIR == 0 ? 1.0 : 0.0;
    constexpr double n_y = DIR == 0 ? 0.0 : 1.0;
    constexpr double tz = DIR == 0 ? 0.0 : -0.0;
    const Gas GL = gas_of(FA[0], T), GR = gas_of(FA[2], T);
    const double u = L.u * n_x + L.v * n_y;
    const double u_R = R.u * n_x + R.v * n_y;
    const double d_u = L.d_u * n_x + L.d_v * n_y;
    const double d_u_R = R.d_u * n_x + R.d_v * n_y;
    constexpr double t_u = tz * n_x + tz * n_y;
    constexpr double t_v = -tz * n_y + tz * n_x;
    L.v = -L.u * n_y + L.v * n_x;
    R.v = -R.u * n_y + R.v * n_x;
    L.d_v = -L.d_u * n_y + L.d_v * n_x;
    R.d_v = -R.d_u * n_y + R.d_v * n_x;
    L.u = u; R.u = u_R; L.d_u = d_u; R.d_u = d_u_R;
    const double Tz[4] = {tz, t_u, t_v, tz};

    double dire[4], mid[4], fo[2];
    grp_q1d<true, CONV>(dire, mid, L, R, Tz, Tz, GL, GR, eps, tol, p_guess, FA, fo);
    const int bad = star_dire_bad(mid, dire, eps);
    if (bad) {
        F[0] = F[1] = F[2] = F[3] = 0.0;
        I[0] = I[1] = I[2] = I[3] = 0.0;
        FZ[0] = FZ[1] = FZ[2] = 0.0;
        return bad;
    }
    const double rho_mid = mid[0] + 0.5 * tau * dire[0];
    const double u_mid = (mid[1] + 0.5 * tau * dire[1]) * n_x - (mid[2] + 0.5 * tau * dire[2]) * n_y;
    const double v_mid = (mid[1] + 0.5 * tau * dire[1]) * n_y + (mid[2] + 0.5 * tau * dire[2]) * n_x;
    const double p_mid = mid[3] + 0.5 * tau * dire[3];
    const double a_mid = fo[0] + 0.5 * tau * fo[1];
    const double g_gm1 = a_mid == 1.0 ? T.a.g_gm1 : (a_mid == 0.0 ? T.b.g_gm1 : 0.0);
    double ge;
    if (a_mid == 1.0 || a_mid == 0.0) {
        ge = g_gm1;
    } else {
        const double gamma_mid = mix_gamma(a_mid, T);
        ge = gamma_mid / (gamma_mid - 1.0);
    }
    F[0] = rho_mid * (u_mid * n_x + v_mid * n_y);
    F[1] = F[0] * u_mid + p_mid * n_x;
    F[2] = F[0] * v_mid + p_mid * n_y;
    F[3] = ge * p_mid / rho_mid + 0.5 * (u_mid * u_mid + v_mid * v_mid);
    F[3] = F[0] * F[3];
    I[1] = (mid[1] + tau * dire[1]) * n_x - (mid[2] + tau * dire[2]) * n_y;
    I[2] = (mid[1] + tau * dire[1]) * n_y + (mid[2] + tau * dire[2]) * n_x;
    I[0] = mid[0] + tau * dire[0];
    I[3] = mid[3] + tau * dire[3];
    const double un_mid = u_mid * n_x + v_mid * n_y;
    FZ[0] = un_mid * a_mid;
    FZ[1] = un_mid;
    FZ[2] = fo[0] + tau * fo[1];
    return 0;
}

// ifvar_check (dim 2) for the split path: tangential derivatives are zero.
__device__ __forceinline__ bool ifvar_bad(const Side &L, const Side &R, double eps)
{
    if (L.p < eps || R.p < eps || L.rho < eps || R.rho < eps) return true;
    return !isfinite(L.d_p) || !isfinite(R.d_p) || !isfinite(L.d_u) || !isfinite(R.d_u)
        || !isfinite(L.d_v) || !isfinite(R.d_v) || !isfinite(L.d_rho) || !isfinite(R.d_rho);
}
// the same with the transversal derivatives checked too (fluid_var_check.c:42-47)
__device__ __forceinline__ bool ifvar_bad_t(const Side &L, const Side &R, const double TL[4], const double TR[4],
                                            double eps)
{
    if (ifvar_bad(L, R, eps)) return true;
    return !isfinite(TL[3]) || !isfinite(TR[3]) || !isfinite(TL[1]) || !isfinite(TR[1])
        || !isfinite(TL[2]) || !isfinite(TR[2]) || !isfinite(TL[0]) || !isfinite(TR[0]);
}

}  // namespace hgrp
