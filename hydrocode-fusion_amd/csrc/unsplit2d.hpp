// unsplit2d.hpp -- one fused launch per time step of the genuinely 2D GRP
// scheme (no dimension splitting).
//
// One launch performs everything the reference's unsplit step spreads over
// bound_cond_slope_limiter_x + bound_cond_slope_limiter_y + flux_generator_x +
// flux_generator_y (Transversal = true) + the update loop of
// GRP_solver_2D_EUL_source (src/finite_volume/grp_solver_2D_EUL_source.c:209-262):
// ghost values, both minmod limiters with the ghost-slope rules across
// directions, x- and y-interface GRP solves over tau with the transversal
// derivatives of the cells beside each interface, the combined conservative
// update and the refresh of both slope sets from the interface values at
// t + tau.  State and slopes are read from one buffer set and written to the
// other (ping-pong), as in the split sweeps.
#pragma once
#include "sweep2d.hpp"

namespace hgrp {

struct UnsplitArgs {
    // Field planes lie `plane` doubles apart, so one address per group reaches
    // every field (fewer scalar registers held across the solves)
    const double *st;                         // state in: rho u v p e
    double *st_o;                             // state out, same layout
    const double *s, *t;                      // x- and y-slopes (rho u v p) of the previous step
    double *s_o, *t_o;                        // refreshed slopes, out
    size_t plane;                             // doubles from one field plane to the next
    const double *bx_lo, *bx_hi;              // BC -1 ghosts of the x-ends from the initial data: [i*4 + q]
    const double *by_lo, *by_hi;              // ... of the y-ends: [j*4 + q]
    int m, n, ld;                             // cells in x, in y, row pitch
    int bc_x, bc_y;                           // config[17], config[18]
    int lim3_x, lim3_y;                       // 0: minmod2 (first step), 1: minmod3 with the previous slope
    int conv;                                 // Newton stop: 0 the reference's (gap <= tol), 1 the converge rule
    int g2d;                                  // interface solver: 0 Q1D (the reference's), 1 G2D (its alternative)
    int form;                                 // launch form: 0 by grid size, 1 tile, 2 strip
    double hh_x, hh_y;                        // h/2
    double tau, nu, mu, eps, alpha, tol;      // tau, tau/h_x, tau/h_y, ...
    RecipDiv div_hx, div_hy;                  // division by h_x, h_y
    Gas gas;
    const MlpPacked *mlp;                     // learned p* guess, or nullptr (the reference's guess)
    SweepStatus *status;                      // [0]: x-interfaces and the update, [1]: y-interfaces
    const SweepStatus *prev;                  // the previous step's pair (nullptr for the first)
};

}  // namespace hgrp
