// sweep2d.hpp -- fused directional sweeps of the dimension-split 2D GRP step.
//
// One launch performs, for one direction, everything the reference spreads over
// bound_cond_slope_limiter_{x,y} + flux_generator_{x,y} + the update loop of
// GRP_solver_2D_split_EUL_source (src/finite_volume/grp_solver_2D_split_EUL_source.c:214-303):
// ghost values, minmod slope limiting, interface reconstruction, the GRP solve,
// the conservative update and the slope refresh from the t+tau interface
// values.  State is read from one buffer set and written to another (ping-pong),
// so neighbouring tiles never see half-updated cells and an aborted sweep
// leaves its input intact.
//
// Grid layout in HBM: SoA, cell (j, i) at j*ld + i, i (y index) contiguous,
// ld >= n a multiple of 32 doubles.  Fields: rho u v p e, x-slopes sx[4],
// y-slopes sy[4] (rho u v p).
#pragma once
#include <stdint.h>

#include "grp_device.hpp"
#include "mlp_guess.hpp"

namespace hgrp {

// boundary condition codes (config[17], config[18]; doc/config.csv)
enum { BC_INITIAL = -1, BC_REFLECT = -2, BC_FREE = -4, BC_PERIODIC = -7, BC_REFLECT_FREE = -24 };

struct SweepStatus {
    unsigned long long fail_key;   // max over failing interfaces of ((loop index << 2) | code) + 1
    unsigned int flags;            // bit 0: positivity failure in the update, bit 1: skipped (halted)
    unsigned int pad;
};

struct SweepArgs {
    // Field planes lie `plane` doubles apart, so one address per group reaches
    // every field (fewer scalar registers held across the solve)
    const double *st;                         // state in: rho u v p e
    double *st_o;                             // state out, same layout
    const double *s;                          // slopes of this direction (rho u v p), in
    double *s_o;                              // refreshed slopes, out
    size_t plane;                             // doubles from one field plane to the next
    const double *bnd_lo, *bnd_hi;            // BC -1 ghost values from the initial data: [line*4 + q]
    int m, n, ld;                             // cells in x, in y, row pitch
    int bc;                                   // boundary condition along the sweep
    int limiter3;                             // 0: minmod2 (first sweep of this direction), 1: minmod3
    int conv;                                 // Newton stop: 0 the reference's (gap <= tol), 1 the converge rule
                                              // (the launchers pick the kernel's Newton mode from it and mlp)
    double h, hh, tau, nu, eps, alpha;        // h, h/2, flux time, update factor, ...
    double tol;                               // Newton tolerance (the reference: eps)
    RecipDiv div_h;                           // division by h
    const MlpPacked *mlp;                     // learned p* guess, or nullptr (the reference's guess)
    Gas gas;                                  // gamma and its derived constants
    // two fluids (sweep2d_tf.hip only): volume fraction of fluid a and its slope
    const double *fa;                         // in
    double *fa_o;                             // out
    const double *sa;                         // slope of this direction, in
    double *sa_o;                             // refreshed slope, out
    TwoGas two;                               // fluids a (config[6]) and b (config[106])
    SweepStatus *status;                      // this sweep
    const SweepStatus *prev;                  // the sweep before (nullptr for the first)
};

// Newton modes of the sweep kernels, a template argument so that each mode's
// solve is compiled on its own (a run-time switch costs registers in all of
// them): the reference's solve; the converge rule (HGRP_NEWTON_CONVERGE); the
// converge rule started from the learned guess.
enum { NM_REF = 0, NM_CONV = 1, NM_LEARNED = 2 };
inline int newton_mode(const SweepArgs &a) { return a.mlp ? NM_LEARNED : (a.conv ? NM_CONV : NM_REF); }

// The kernel's argument block (its first parameter) read again at the point of
// the call: the copy's fields are loaded there from kernel-argument memory
// (scalar-cache loads) through an address the compiler cannot trace back to
// the loads at the top of the kernel, so values used only after the solves
// are not held in scalar registers across them.
template <class Args>
__device__ __forceinline__ Args reload_args()
{
#if defined(__HIP_DEVICE_COMPILE__)
    using P = __attribute__((address_space(4))) const Args *;
    P p = (P)__builtin_amdgcn_kernarg_segment_ptr();
    __asm__ volatile("" : "+s"(p));
    return *p;
#else
    return Args{};   // device code only
#endif
}

// Interface state from a cell value w and its slope s along the interface
// normal: w + h/2 s on the left (lower) side, w - h/2 s on the right (upper)
// side, with the slope as the normal derivative (flux_generator_x.c:45-86).
__device__ __forceinline__ Side recon(const double w[4], const double s[4], double hh, bool left)
{
    Side f;
    if (left) {
        f.rho = w[0] + hh * s[0]; f.u = w[1] + hh * s[1]; f.v = w[2] + hh * s[2]; f.p = w[3] + hh * s[3];
    } else {
        f.rho = w[0] - hh * s[0]; f.u = w[1] - hh * s[1]; f.v = w[2] - hh * s[2]; f.p = w[3] - hh * s[3];
    }
    f.d_rho = s[0]; f.d_u = s[1]; f.d_v = s[2]; f.d_p = s[3];
    return f;
}

}  // namespace hgrp
