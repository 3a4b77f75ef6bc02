// Host build of the device pow (hydrocode-fusion_amd/csrc/fastpow.hpp) so its
// rounding can be checked on the CPU against glibc and mpmath.
#include "fastpow.hpp"

extern "C" void hg_fast_pow_many(const double *x, const double *y, int n, double *out)
{
    for (int k = 0; k < n; ++k) out[k] = hgrp::fast_pow(x[k], y[k]);
}

extern "C" void hg_cube_many(const double *x, int n, double *out)
{
    for (int k = 0; k < n; ++k) out[k] = hgrp::cube(x[k]);
}

extern "C" void hg_log_dd(double x, double *hi, double *lo)
{
    const hgrp::DD d = hgrp::log_dd(x);
    *hi = d.hi;
    *lo = d.lo;
}

extern "C" double hg_exp_dd(double hi, double lo)
{
    return hgrp::exp_dd(hgrp::DD{hi, lo});
}
