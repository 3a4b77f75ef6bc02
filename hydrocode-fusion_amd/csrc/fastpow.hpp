// fastpow.hpp -- FP64 pow for positive normal bases, built from a double-double
// natural log and a double-double exp (tables in pow_tables.hpp).
//
// Error before the final rounding is about 2^-70 relative (2^-18 ulp), so the
// result is the correctly rounded x^y except when x^y lies within that
// distance of a rounding boundary; it therefore agrees with any other
// near-correctly-rounded libm (glibc's pow among them) on all but a tiny
// fraction of arguments, which keeps GPU and CPU runs of the GRP solver in
// step.  It is also several times cheaper than a general-purpose pow, and the
// log can be shared between powers of the same base (PowBase).
//
// Method (textbook range reduction, e.g. Muller, "Elementary Functions", ch. 6-7):
//   log: x = 2^k m, m in [1,2); r = m*INVC[j] - 1 exact (|r| < 2^-7);
//        (the tables are one array, pow_tables.hpp, so a kernel keeps one address)
//        log x = k ln2 + LOGC[j] + log1p(r), log1p by its Taylor series to r^10.
//   exp: y log x = (k2/128) ln2 + z, |z| <= ln2/256;
//        exp = 2^(k2 >> 7) * EXP2[k2 & 127] * (1 + z + ... + z^7/7!).
// Bases outside (0, inf) normal, and results near overflow/underflow, take the
// platform pow so that special values keep their C semantics.
#pragma once
#include <math.h>

#if defined(__HIP__) || defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HGRP_HD __host__ __device__ __forceinline__
#define HGRP_COLD static __host__ __device__ __attribute__((noinline))
#define HGRP_POW_TABLE __device__ __constant__ static
#else
#define HGRP_HD static inline
#define HGRP_COLD static __attribute__((noinline))
#define HGRP_POW_TABLE static
#endif

#include "pow_tables.hpp"

namespace hgrp {

struct DD {
    double hi, lo;
};

HGRP_HD DD two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

// log(x) as a double-double, x positive normal finite
HGRP_HD DD log_dd(double x)
{
    int k;
    double m = frexp(x, &k);           // m in [0.5, 1)
    m *= 2.0;
    k -= 1;                            // x = 2^k m, m in [1, 2)
    const int j = (int)((m - 1.0) * 128.0);
    const double *lt = powtab::TAB + powtab::LOG_ROW * j;
    const double invc = lt[0];
    const double r = fma(m, invc, -1.0);                  // exact
    const double q_hi = r * r;
    const double q_lo = fma(r, r, -q_hi);                  // r^2 = q_hi + q_lo exactly
    // log1p(r) - r + r^2/2 = r^3 (1/3 - r/4 + ... + r^7/10)
    double t = -1.0 / 10.0;
    t = fma(t, r, 1.0 / 9.0);
    t = fma(t, r, -1.0 / 8.0);
    t = fma(t, r, 1.0 / 7.0);
    t = fma(t, r, -1.0 / 6.0);
    t = fma(t, r, 1.0 / 5.0);
    t = fma(t, r, -1.0 / 4.0);
    t = fma(t, r, 1.0 / 3.0);
    const double tail = t * (q_hi * r);
    const double kd = (double)k;
    const DD s1 = two_sum(kd * powtab::LN2_HI, lt[1]);
    const DD s2 = two_sum(s1.hi, r);
    const DD s3 = two_sum(s2.hi, -0.5 * q_hi);
    const double lo = s1.lo + s2.lo + s3.lo + kd * powtab::LN2_LO + lt[2] - 0.5 * q_lo + tail;
    const double hi = s3.hi + lo;
    return {hi, lo - (hi - s3.hi)};
}

// exp(p.hi + p.lo), |p.hi| < 700
HGRP_HD double exp_dd(DD p)
{
    const double kf = rint(p.hi * powtab::INV_LN2N);
    const int k2 = (int)kf;
    // z = p - kf ln2/128 renormalised so that |z_lo| <= 2^-53 |z_hi| (the
    // polynomial below is evaluated at z_hi)
    const DD z = two_sum(p.hi - kf * powtab::LN2N_HI, p.lo - kf * powtab::LN2N_LO);
    const double z_hi = z.hi, z_lo = z.lo;
    // exp(z) - 1 - z_hi = z_lo + z_hi z_lo + z_hi^2 (1/2 + z_hi/6 + ... + z_hi^5/7!)
    double t = 1.0 / 5040.0;
    t = fma(t, z_hi, 1.0 / 720.0);
    t = fma(t, z_hi, 1.0 / 120.0);
    t = fma(t, z_hi, 1.0 / 24.0);
    t = fma(t, z_hi, 1.0 / 6.0);
    t = fma(t, z_hi, 0.5);
    const double small = z_lo + z_hi * z_lo + (z_hi * z_hi) * t;
    const int j = k2 & 127;
    const double *et = powtab::TAB + powtab::EXP_BASE + powtab::EXP_ROW * j;
    const double T_hi = et[0], T_lo = et[1];
    const double prod_hi = T_hi * z_hi;
    const double prod_lo = fma(T_hi, z_hi, -prod_hi);
    const DD s = two_sum(T_hi, prod_hi);
    const double tail = s.lo + prod_lo + T_hi * small + T_lo + T_lo * (z_hi + small);
    return ldexp(s.hi + tail, k2 >> 7);
}

// The platform pow for what the fast path leaves (bases that are not positive
// normal numbers, results near overflow or underflow).  Out of line: pow_with
// is inlined at dozens of sites of the sweep kernels, and a copy of the
// library pow at each would double their code size for a path they rarely take.
HGRP_COLD double pow_special(double x, double y) { return pow(x, y); }

HGRP_HD bool pow_base_ok(double x) { return x >= 2.2250738585072014e-308 && x < INFINITY; }

// A base whose log is computed once and raised to several exponents.
struct PowBase {
    double x;
    DD L;
    bool ok;
};

HGRP_HD PowBase pow_base(double x)
{
    PowBase b;
    b.x = x;
    b.ok = pow_base_ok(x);
    b.L = b.ok ? log_dd(x) : DD{0.0, 0.0};
    return b;
}

HGRP_HD double pow_with(const PowBase &b, double y)
{
    if (y == 0.0 || b.x == 1.0) return 1.0;
    if (b.ok) {
        const double p_hi = y * b.L.hi;
        if (fabs(p_hi) < 700.0) {
            const double p_lo = fma(y, b.L.hi, -p_hi) + y * b.L.lo;
            return exp_dd(two_sum(p_hi, p_lo));
        }
    }
    return pow_special(b.x, y);
}

HGRP_HD double fast_pow(double x, double y) { return pow_with(pow_base(x), y); }

// x^3 rounded once: x^2 and x^3 formed exactly as double-doubles
HGRP_HD double cube(double x)
{
    const double s = x * x;
    const double s_lo = fma(x, x, -s);
    const double p = s * x;
    const double p_lo = fma(s, x, -p);
    return p + (p_lo + s_lo * x);
}

}  // namespace hgrp
