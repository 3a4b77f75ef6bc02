// unsplit2d.hip -- the genuinely 2D GRP step (no dimension splitting), one
// launch per time step.
//
//   unsplit_step_kernel  a block of W waves (W = TILE_W) covers 62 consecutive y-cells
//                        (lanes: cells i0-1 .. i0+62) of the cell rows
//                        j0-1 .. j0+W-2 (wave w: row j0-1+w).  Every lane forms
//                        the "side data" of its cell -- value, limited x-slope,
//                        limited y-slope, or the ghost's value and slopes at the
//                        edges -- and parks it in LDS.  Waves 0..W-2 solve the
//                        x-interface between their row and the next; waves
//                        1..W-2 also solve the y-interfaces of the updated rows
//                        (each half a pair of rows, 2 x 32 interfaces) and
//                        update their 62 cells with both flux differences.
//                        Rows j0-1 and j0+W-2 are halo rows; interface rows and
//                        columns on tile edges are solved by both tiles,
//                        everything else once.
//
// Semantics follow src/finite_volume/grp_solver_2D_EUL_source.c:181-262,
// src/inter_process/bound_cond_slope_limiter_{x,y}.c:27-129 (including the
// ghost slopes across directions, :111-126) and
// src/flux_calc/flux_generator_{x,y}.c:31-165 with Transversal = true.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include "grp_device.hpp"
#include "g2d_device.hpp"
#include "unsplit2d.hpp"

namespace hgrp {

namespace {

// Values, limited slopes and interface side data around one step.  Rows c
// (x index) run over [-1, m], columns ci (y index) over [-1, n]; -1, m and n
// are ghost cells.
struct Grid {
    const UnsplitArgs &a;
    __device__ explicit Grid(const UnsplitArgs &args) : a(args) {}

    __device__ __forceinline__ size_t at(int c, int ci) const { return (size_t)c * (size_t)a.ld + (size_t)ci; }
    __device__ __forceinline__ void cell(int c, int ci, double w[4]) const
    {
        const size_t o = at(c, ci);
        const double *f = a.st + o;
        w[0] = f[0]; w[1] = f[a.plane]; w[2] = f[2 * a.plane]; w[3] = f[3 * a.plane];
    }
    // value of row c in [-1, m] of column ci in [0, n): x-ghosts per config[17]
    // (bound_cond_slope_limiter_x.c:34-82)
    __device__ __forceinline__ void xval(int c, int ci, double w[4]) const
    {
        if (c >= 0 && c < a.m) { cell(c, ci, w); return; }
        const bool lo = c < 0;
        switch (a.bc_x) {
        case BC_INITIAL: {
            const double *b = (lo ? a.bx_lo : a.bx_hi) + (size_t)ci * 4;
            w[0] = b[0]; w[1] = b[1]; w[2] = b[2]; w[3] = b[3];
            return;
        }
        case BC_PERIODIC: cell(lo ? a.m - 1 : 0, ci, w); return;
        case BC_REFLECT: cell(lo ? 0 : a.m - 1, ci, w); w[1] = -w[1]; return;
        case BC_REFLECT_FREE: cell(lo ? 0 : a.m - 1, ci, w); if (lo) w[1] = -w[1]; return;
        default: cell(lo ? 0 : a.m - 1, ci, w); return;
        }
    }
    // value of column ci in [-1, n] of row c in [0, m): y-ghosts per config[18]
    // (bound_cond_slope_limiter_y.c:34-82)
    __device__ __forceinline__ void yval(int c, int ci, double w[4]) const
    {
        if (ci >= 0 && ci < a.n) { cell(c, ci, w); return; }
        const bool lo = ci < 0;
        switch (a.bc_y) {
        case BC_INITIAL: {
            const double *b = (lo ? a.by_lo : a.by_hi) + (size_t)c * 4;
            w[0] = b[0]; w[1] = b[1]; w[2] = b[2]; w[3] = b[3];
            return;
        }
        case BC_PERIODIC: cell(c, lo ? a.n - 1 : 0, w); return;
        case BC_REFLECT: cell(c, lo ? 0 : a.n - 1, w); w[2] = -w[2]; return;
        case BC_REFLECT_FREE: cell(c, lo ? 0 : a.n - 1, w); if (lo) w[2] = -w[2]; return;
        default: cell(c, lo ? 0 : a.n - 1, w); return;
        }
    }
    // minmod slope from the neighbours (slope_limiter_2D_x.c:32-84, slope_limiter.c:31-83)
    __device__ __forceinline__ void limit(const double vm[4], const double v0[4], const double vp[4], const double *prev,
                          size_t o, int lim3, const RecipDiv &d, double s[4]) const
    {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double s_L = div_by(v0[q] - vm[q], d);
            const double s_R = div_by(vp[q] - v0[q], d);
            s[q] = lim3 ? minmod3(a.alpha * s_L, a.alpha * s_R, prev[o + q * a.plane]) : minmod2(s_L, s_R);
        }
    }
    __device__ __forceinline__ void lim_x(int c, int ci, const double v0[4], double s[4]) const
    {
        double vm[4], vp[4];
        xval(c - 1, ci, vm);
        xval(c + 1, ci, vp);
        limit(vm, v0, vp, a.s, at(c, ci), a.lim3_x, a.div_hx, s);
    }
    __device__ __forceinline__ void lim_y(int c, int ci, const double v0[4], double t[4]) const
    {
        double vm[4], vp[4];
        yval(c, ci - 1, vm);
        yval(c, ci + 1, vp);
        limit(vm, v0, vp, a.t, at(c, ci), a.lim3_y, a.div_hy, t);
    }
    __device__ __forceinline__ void lim_x_at(int c, int ci, double s[4]) const
    {
        double v0[4];
        cell(c, ci, v0);
        lim_x(c, ci, v0, s);
    }
    __device__ __forceinline__ void lim_y_at(int c, int ci, double t[4]) const
    {
        double v0[4];
        cell(c, ci, v0);
        lim_y(c, ci, v0, t);
    }

    // Side data of (c, ci): value v, x-slope s, y-slope t.  A cell: its limited
    // slopes.  An x-ghost (c = -1 or m): bfv_L/R -- ghost value, the x-slope
    // rule of bound_cond_slope_limiter_x.c:94-109 and, across, the y-slopes of
    // the boundary cell (bound_cond_slope_limiter_y.c:111-126).  A y-ghost
    // (ci = -1 or n): bfv_D/U with the mirrored rules.  Slopes the reference
    // never sets stay zero (its ghost arrays are calloc'ed).  Not for corners.
    __device__ __forceinline__ void side(int c, int ci, double v[4], double s[4], double t[4]) const
    {
        const bool xin = c >= 0 && c < a.m, yin = ci >= 0 && ci < a.n;
#pragma unroll
        for (int q = 0; q < 4; ++q) { s[q] = 0.0; t[q] = 0.0; }
        if (xin && yin) {
            cell(c, ci, v);
            lim_x(c, ci, v, s);
            lim_y(c, ci, v, t);
        } else if (yin) {
            const bool lo = c < 0;
            const int near = lo ? 0 : a.m - 1, far = lo ? a.m - 1 : 0;
            xval(c, ci, v);
            double q4[4];
            switch (a.bc_x) {
            case BC_REFLECT: lim_x_at(near, ci, q4); s[1] = q4[1]; break;
            case BC_PERIODIC: lim_x_at(far, ci, s); break;
            case BC_REFLECT_FREE: if (lo) { lim_x_at(0, ci, q4); s[1] = q4[1]; } break;
            default: break;
            }
            switch (a.bc_x) {
            case BC_REFLECT: case BC_FREE: case BC_REFLECT_FREE: lim_y_at(near, ci, t); break;
            case BC_PERIODIC: lim_y_at(far, ci, t); break;
            default: break;
            }
        } else {
            const bool lo = ci < 0;
            const int near = lo ? 0 : a.n - 1, far = lo ? a.n - 1 : 0;
            yval(c, ci, v);
            double q4[4];
            switch (a.bc_y) {
            case BC_REFLECT: lim_y_at(c, near, q4); t[2] = q4[2]; break;
            case BC_PERIODIC: lim_y_at(c, far, t); break;
            case BC_REFLECT_FREE: if (lo) { lim_y_at(c, 0, q4); t[2] = q4[2]; } break;
            default: break;
            }
            switch (a.bc_y) {
            case BC_REFLECT: case BC_FREE: case BC_REFLECT_FREE: lim_x_at(c, near, s); break;
            case BC_PERIODIC: lim_x_at(c, far, s); break;
            default: break;
            }
        }
    }
};

// GRP solve of one interface with transversal derivatives, normal x (dir 0)
// or y (dir 1); failures are recorded in status[dir] as
// ((loop index << 2) | code) + 1, the reference's loop order
// G2: GRP_2D_flux with its alternative solver, G2D (grp_2d_flux_g2d).
template <int NM, bool G2>
__device__ __forceinline__ void solve_t(const UnsplitArgs &a, int dir, const Side &SL, const Side &SR,
                                        const double TL[4], const double TR[4], unsigned long long loop_index,
                                        double F[4], double I[4], double p_guess)
{
    int code = ifvar_bad_t(SL, SR, TL, TR, a.eps) ? 1 : 0;
    const double n_x = dir ? 0.0 : 1.0, n_y = dir ? 1.0 : 0.0;
    const int bad = G2 ? grp_2d_flux_g2d<NM != NM_REF>(SL, SR, TL, TR, a.gas, n_x, n_y, a.tau, a.eps, F, I, a.tol,
                                                       p_guess)
                       : grp_2d_flux_n<NM != NM_REF>(SL, SR, TL, TR, a.gas, n_x, n_y, a.tau, a.eps, F, I, a.tol,
                                                     p_guess);
    if (bad) code = 2;
    if (code) atomicMax(&a.status[dir].fail_key, ((loop_index << 2) | (unsigned long long)code) + 1ull);
}

// Learned starting pressure for the interface (SL | SR) with normal x (dir 0)
// or y (dir 1), every lane of the wave (uniform call); -1 outside NM_LEARNED,
// where Newton will not run (the acoustic shortcut, the equal-state exit) or
// where the lane has no interface.  The states are those normal to the
// interface (GRP_2D_flux's rotation keeps rho and p; u_n is u or v).
template <int NM>
__device__ __forceinline__ double unsplit_guess(const UnsplitArgs &a, bool active, int dir, const Side &SL,
                                                const Side &SR, uint32_t *lds)
{
    if constexpr (NM != NM_LEARNED) {
        return -1.0;
    } else {
        const double uL = dir ? SL.v : SL.u, uR = dir ? SR.v : SR.u;
        const double dist = sqrt((SL.rho - SR.rho) * (SL.rho - SR.rho) + (uL - uR) * (uL - uR)
                                 + (SL.p - SR.p) * (SL.p - SR.p));
        const bool want = active && !(dist < a.eps) && !(fabs(uL - uR) < a.tol && fabs(SL.p - SR.p) < a.tol);
        return learned_guess_wave(a.mlp, want, SL.rho, uL, SL.p, SR.rho, uR, SR.p, a.gas.g, a.gas.g, lds);
    }
}

}  // namespace

// TRIM: the y-interface results go to the slope slots of the side data (dead
// after both solve passes; one more barrier) instead of an array of their own,
// where the block's LDS would not fit otherwise.
template <int W, int NM, bool G2, bool TRIM>
__device__ __forceinline__ void unsplit_step_body(const UnsplitArgs &a)
{
    __shared__ double sd[W][12][64];     // side data of row c: value, x-slope, y-slope
    __shared__ double fx[W - 1][8][64];  // x-interface c+1: flux, values at t + tau
    __shared__ double gy[TRIM ? 1 : W - 2][8][64];  // y-interfaces of the updated rows: flux, values at t + tau
    __shared__ uint32_t mlp_lds[NM == NM_LEARNED ? W : 1][NM == NM_LEARNED ? 64 * 8 : 1];   // network inputs
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = a.m, n = a.n;
    const int i0 = blockIdx.x * 62, ci = i0 - 1 + lane;
    const int j0 = blockIdx.y * (W - 2), c = j0 - 1 + w;
    const Grid G(a);
    const bool xin = c >= 0 && c < m, yin = ci >= 0 && ci < n;
    {
        double v[4] = {1.0, 0.0, 0.0, 1.0}, s[4] = {0.0, 0.0, 0.0, 0.0}, t[4] = {0.0, 0.0, 0.0, 0.0};
        if (c <= m && ci <= n && (xin || yin)) G.side(c, ci, v, s, t);   // a cell or a ghost, not a corner
#pragma unroll
        for (int q = 0; q < 4; ++q) { sd[w][q][lane] = v[q]; sd[w][4 + q][lane] = s[q]; sd[w][8 + q][lane] = t[q]; }
    }
    // a failed step ends the run: later steps do nothing.  The status is read
    // after the side loads went out (they stay in bounds either way); the
    // whole block returns together.
    if (a.prev != nullptr) {
        const SweepStatus p0 = a.prev[0], p1 = a.prev[1];
        if (p0.fail_key != 0ull || p0.flags != 0u || p1.fail_key != 0ull) {
            if (threadIdx.x == 0) atomicOr(&a.status[0].flags, 2u);
            return;
        }
    }
    __syncthreads();

    // pass 0: x-interface f = c + 1 between rows c and c + 1 (flux_generator_x.c:42-121),
    //         the right side from the next wave;
    // pass 1: y-interfaces of the updated rows (flux_generator_y.c:42-121), normal derivative =
    //         y-slope, transversal = -(x-slope); wave w takes half a row pair (2 rows x 32
    //         interfaces), so a front lying along y meets half as many waves as with one
    //         64-lane row per wave.
    // One solve site for both, so the GRP code and its registers are laid out once.
    const int f = c + 1;
    const int task = w - 1;                                  // waves 1 .. W-2
    const int yr = 1 + 2 * (task >> 1) + (lane >> 5);        // tile row of this lane's y-interface
    const int yt = ((task & 1) << 5) + (lane & 31);          // between tile lanes yt and yt + 1
    const int yc = j0 - 1 + yr, ycell = i0 - 1 + yt;
    double Fy[4] = {0.0, 0.0, 0.0, 0.0}, Iyv[4] = {0.0, 0.0, 0.0, 0.0};   // TRIM: pass 1's results
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        Side SL, SR;
        double TL[4], TR[4];
        unsigned long long idx = 0;
        bool active;
        if (pass == 0) {
            active = w <= W - 2 && lane >= 1 && lane <= 62 && yin && f <= m;
            if (active) {
                double vl[4], sl[4], vr[4], sr[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    vl[q] = sd[w][q][lane];     sl[q] = sd[w][4 + q][lane];     TL[q] = sd[w][8 + q][lane];
                    vr[q] = sd[w + 1][q][lane]; sr[q] = sd[w + 1][4 + q][lane]; TR[q] = sd[w + 1][8 + q][lane];
                }
                SL = recon(vl, sl, a.hh_x, true);
                SR = recon(vr, sr, a.hh_x, false);
                idx = (unsigned long long)ci * (unsigned long long)(m + 1) + (unsigned long long)f;
            }
        } else {
            active = w >= 1 && w <= W - 2 && yc >= 0 && yc < m && yt <= 62 && ycell <= n - 1;
            if (active) {
                double vd[4], td[4], vu[4], tu[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    vd[q] = sd[yr][q][yt];     td[q] = sd[yr][8 + q][yt];     TL[q] = -sd[yr][4 + q][yt];
                    vu[q] = sd[yr][q][yt + 1]; tu[q] = sd[yr][8 + q][yt + 1]; TR[q] = -sd[yr][4 + q][yt + 1];
                }
                SL = recon(vd, td, a.hh_y, true);
                SR = recon(vu, tu, a.hh_y, false);
                idx = (unsigned long long)yc * (unsigned long long)(n + 1) + (unsigned long long)(ycell + 1);
            }
        }
        double F[4] = {0.0, 0.0, 0.0, 0.0}, I[4] = {0.0, 0.0, 0.0, 0.0};
        const double pg = unsplit_guess<NM>(a, active, pass, SL, SR, mlp_lds[NM == NM_LEARNED ? w : 0]);
        // the argument block is read again for the solve, so no scalar register holds it across
        if (active) solve_t<NM, G2>(reload_args<UnsplitArgs>(), pass, SL, SR, TL, TR, idx, F, I, pg);
        if (pass == 0 && w <= W - 2) {
#pragma unroll
            for (int q = 0; q < 4; ++q) { fx[w][q][lane] = F[q]; fx[w][4 + q][lane] = I[q]; }
        }
        if (pass == 1 && w >= 1 && w <= W - 2) {
            if constexpr (TRIM) {
#pragma unroll
                for (int q = 0; q < 4; ++q) { Fy[q] = F[q]; Iyv[q] = I[q]; }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) { gy[yr - 1][q][yt] = F[q]; gy[yr - 1][4 + q][yt] = I[q]; }
            }
        }
    }
    __syncthreads();
    if constexpr (TRIM) {
        // every wave is past both passes: the slopes of the side data are dead
        if (w >= 1 && w <= W - 2) {
#pragma unroll
            for (int q = 0; q < 4; ++q) { sd[yr][4 + q][yt] = Fy[q]; sd[yr][8 + q][yt] = Iyv[q]; }
        }
        __syncthreads();
    }

    // combined update and slope refresh (grp_solver_2D_EUL_source.c:231-262)
    if (w >= 1 && w <= W - 2 && xin && yin && lane >= 1 && lane <= 62) {
        const UnsplitArgs a = reload_args<UnsplitArgs>();
        const size_t o = G.at(c, ci);
        const double e_c = a.st[o + 4 * a.plane];
        double vc[4], Fl[4], Ixl[4], Fr[4], Ixr[4], Gy[4], Iy[4], Gl[4], Iyl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            vc[q] = sd[w][q][lane];
            Fl[q] = fx[w - 1][q][lane]; Ixl[q] = fx[w - 1][4 + q][lane];
            Fr[q] = fx[w][q][lane];     Ixr[q] = fx[w][4 + q][lane];
            if constexpr (TRIM) {
                Gl[q] = sd[w][4 + q][lane - 1]; Iyl[q] = sd[w][8 + q][lane - 1];
                Gy[q] = sd[w][4 + q][lane];     Iy[q] = sd[w][8 + q][lane];
            } else {
                Gl[q] = gy[w - 1][q][lane - 1]; Iyl[q] = gy[w - 1][4 + q][lane - 1];
                Gy[q] = gy[w - 1][q][lane];     Iy[q] = gy[w - 1][4 + q][lane];
            }
        }
        const double nu = a.nu, mu = a.mu;
        const double mom_x = vc[0] * vc[1] - nu * (Fr[1] - Fl[1]) - mu * (Gy[1] - Gl[1]);
        const double mom_y = vc[0] * vc[2] - nu * (Fr[2] - Fl[2]) - mu * (Gy[2] - Gl[2]);
        const double ene = vc[0] * e_c - nu * (Fr[3] - Fl[3]) - mu * (Gy[3] - Gl[3]);
        const double rho = vc[0] + (-nu * (Fr[0] - Fl[0]) - mu * (Gy[0] - Gl[0]));
        const double uu = mom_x / rho;
        const double vv = mom_y / rho;
        const double p = (ene - 0.5 * mom_x * uu - 0.5 * mom_y * vv) * a.gas.gm1();
        double *fo = a.st_o + o;
        const size_t P = a.plane;
        fo[0] = rho; fo[P] = uu; fo[2 * P] = vv; fo[4 * P] = ene / rho; fo[3 * P] = p;
        if (p < a.eps || rho < a.eps) atomicOr(&a.status[0].flags, 1u);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a.s_o[o + q * P] = div_by(Ixr[q] - Ixl[q], a.div_hx);
            a.t_o[o + q * P] = div_by(Iy[q] - Iyl[q], a.div_hy);
        }
    }
}

// LDS of a tile of W rows: side data, x- and y-interface results, the
// learned guess's network inputs; over 160 KiB the y results move into the
// side data (TRIM)
template <int W, int NM>
constexpr bool tile_trim()
{
    return (size_t)W * 12 * 512 + (size_t)(W - 1) * 8 * 512 + (size_t)(W - 2) * 8 * 512 +
               (NM == NM_LEARNED ? (size_t)W * 2048 : 0) > 160 * 1024;
}

template <int W, int MINW, int NM>
__global__ void __launch_bounds__(64 * W, MINW) unsplit_step_kernel(UnsplitArgs a)
{
    unsplit_step_body<W, NM, false, tile_trim<W, NM>()>(a);
}

// the same step with GRP_2D_flux's alternative interface solver, G2D
template <int W, int MINW, int NM>
__global__ void __launch_bounds__(64 * W, MINW) unsplit_step_g2d_kernel(UnsplitArgs a)
{
    unsplit_step_body<W, NM, true, tile_trim<W, NM>()>(a);
}

// Strip form: one wave per block walks down R cell rows of a 62-cell strip
// (lanes: cells i0-1 .. i0+62 of the row).  Iteration r (from r0-1) forms the
// side data of row r+1, solves the y-interfaces of row r and the x-interface
// r+1 (one solve site, two passes), and updates row r from the x-interface r
// kept from the iteration before.  No wave waits for another: the block is
// one wave, its LDS holds two rows of side data and two rows of interface
// results (16 KiB).  One extra x-interface per R rows.
template <int NM, bool G2>
__device__ __forceinline__ void unsplit_strip_body(const UnsplitArgs &a, int R)
{
    __shared__ double sd[2][12][64];   // side data of rows r and r + 1 (alternating)
    __shared__ double fl[8][64];       // x-interface r: flux, values at t + tau
    __shared__ double gb[8][64];       // y-interfaces of row r: flux, values at t + tau
    __shared__ uint32_t mlp_lds[NM == NM_LEARNED ? 64 * 8 : 1];   // network inputs
    if (a.prev != nullptr) {
        const SweepStatus p0 = a.prev[0], p1 = a.prev[1];
        if (p0.fail_key != 0ull || p0.flags != 0u || p1.fail_key != 0ull) {
            if (threadIdx.x == 0) atomicOr(&a.status[0].flags, 2u);
            return;
        }
    }
    const int lane = threadIdx.x;
    const int m = a.m, n = a.n;
    const int ci = blockIdx.x * 62 - 1 + lane;
    const int r0 = blockIdx.y * R, r1 = min(r0 + R, m);   // rows r0 .. r1-1 are updated here
    const Grid G(a);
    const bool yin = ci >= 0 && ci < n;
    const bool colv = ci <= n;                             // a cell or a y-ghost column

    auto put_side = [&](int buf, int c) {
        double v[4] = {1.0, 0.0, 0.0, 1.0}, s[4] = {0.0, 0.0, 0.0, 0.0}, t[4] = {0.0, 0.0, 0.0, 0.0};
        const UnsplitArgs ar = reload_args<UnsplitArgs>();   // nothing of the loads is held across the solves
        if (colv && c <= m && ((c >= 0 && c < m) || yin)) Grid(ar).side(c, ci, v, s, t);
#pragma unroll
        for (int q = 0; q < 4; ++q) { sd[buf][q][lane] = v[q]; sd[buf][4 + q][lane] = s[q]; sd[buf][8 + q][lane] = t[q]; }
    };
    int cb = 0;
    put_side(cb, r0 - 1);
#pragma unroll 1
    for (int r = r0 - 1; r < r1; ++r) {
        const int nb = 1 - cb;
        put_side(nb, r + 1);
        __syncthreads();
        const bool xin = r >= r0;                          // row r is updated (r0 <= r < r1 <= m)
        double F[4], I[4];
#pragma unroll 1
        for (int pass = 0; pass < 2; ++pass) {
            Side SL, SR;
            double TL[4], TR[4];
            unsigned long long idx = 0;
            bool active;
            if (pass == 0) {   // y-interface ci + 1 of row r: normal = y-slope, transversal = -(x-slope)
                active = xin && lane <= 62 && ci <= n - 1;
                if (active) {
                    double vd[4], td[4], vu[4], tu[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        vd[q] = sd[cb][q][lane];     td[q] = sd[cb][8 + q][lane];     TL[q] = -sd[cb][4 + q][lane];
                        vu[q] = sd[cb][q][lane + 1]; tu[q] = sd[cb][8 + q][lane + 1]; TR[q] = -sd[cb][4 + q][lane + 1];
                    }
                    SL = recon(vd, td, a.hh_y, true);
                    SR = recon(vu, tu, a.hh_y, false);
                    idx = (unsigned long long)r * (unsigned long long)(n + 1) + (unsigned long long)(ci + 1);
                }
            } else {           // x-interface r + 1 between rows r and r + 1
                active = lane >= 1 && lane <= 62 && yin && r + 1 <= m;
                if (active) {
                    double vl[4], sl[4], vr[4], sr[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        vl[q] = sd[cb][q][lane]; sl[q] = sd[cb][4 + q][lane]; TL[q] = sd[cb][8 + q][lane];
                        vr[q] = sd[nb][q][lane]; sr[q] = sd[nb][4 + q][lane]; TR[q] = sd[nb][8 + q][lane];
                    }
                    SL = recon(vl, sl, a.hh_x, true);
                    SR = recon(vr, sr, a.hh_x, false);
                    idx = (unsigned long long)ci * (unsigned long long)(m + 1) + (unsigned long long)(r + 1);
                }
            }
            double Fo[4] = {0.0, 0.0, 0.0, 0.0}, Io[4] = {0.0, 0.0, 0.0, 0.0};
            const double pg = unsplit_guess<NM>(a, active, 1 - pass, SL, SR, mlp_lds);
            if (active) solve_t<NM, G2>(reload_args<UnsplitArgs>(), 1 - pass, SL, SR, TL, TR, idx, Fo, Io, pg);
            if (pass == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) { gb[q][lane] = Fo[q]; gb[4 + q][lane] = Io[q]; }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) { F[q] = Fo[q]; I[q] = Io[q]; }
        }
        __syncthreads();
        if (xin && yin && lane >= 1 && lane <= 62) {
            const UnsplitArgs a = reload_args<UnsplitArgs>();
            const size_t o = G.at(r, ci);
            const double e_c = a.st[o + 4 * a.plane];
            double vc[4], Fl[4], Ixl[4], Gy[4], Gl[4], Iy[4], Iyl[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                vc[q] = sd[cb][q][lane];
                Fl[q] = fl[q][lane]; Ixl[q] = fl[4 + q][lane];
                Gy[q] = gb[q][lane]; Iy[q] = gb[4 + q][lane];
                Gl[q] = gb[q][lane - 1]; Iyl[q] = gb[4 + q][lane - 1];
            }
            const double nu = a.nu, mu = a.mu;
            const double mom_x = vc[0] * vc[1] - nu * (F[1] - Fl[1]) - mu * (Gy[1] - Gl[1]);
            const double mom_y = vc[0] * vc[2] - nu * (F[2] - Fl[2]) - mu * (Gy[2] - Gl[2]);
            const double ene = vc[0] * e_c - nu * (F[3] - Fl[3]) - mu * (Gy[3] - Gl[3]);
            const double rho = vc[0] + (-nu * (F[0] - Fl[0]) - mu * (Gy[0] - Gl[0]));
            const double uu = mom_x / rho;
            const double vv = mom_y / rho;
            const double p = (ene - 0.5 * mom_x * uu - 0.5 * mom_y * vv) * a.gas.gm1();
            double *fo = a.st_o + o;
            const size_t P = a.plane;
            fo[0] = rho; fo[P] = uu; fo[2 * P] = vv; fo[4 * P] = ene / rho; fo[3 * P] = p;
            if (p < a.eps || rho < a.eps) atomicOr(&a.status[0].flags, 1u);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a.s_o[o + q * P] = div_by(I[q] - Ixl[q], a.div_hx);
                a.t_o[o + q * P] = div_by(Iy[q] - Iyl[q], a.div_hy);
            }
        }
        // x-interface r + 1 is the next row's lower interface (each lane keeps its own slot)
#pragma unroll
        for (int q = 0; q < 4; ++q) { fl[q][lane] = F[q]; fl[4 + q][lane] = I[q]; }
        cb = nb;
    }
}

template <int NM>
__global__ void __launch_bounds__(64, 2) unsplit_strip_kernel(UnsplitArgs a, int R)
{
    unsplit_strip_body<NM, false>(a, R);
}

// the same with GRP_2D_flux's alternative interface solver, G2D
template <int NM>
__global__ void __launch_bounds__(64, 2) unsplit_strip_g2d_kernel(UnsplitArgs a, int R)
{
    unsplit_strip_body<NM, true>(a, R);
}

// ---- launchers -----------------------------------------------------------------
// Two forms: the tile (TILE_W waves, TILE_W - 2 updated rows per block) and the
// strip (one wave per block walking down its rows).  The context picks one
// (hgrp2d_set_unsplit_form, hgrp2d_tune) or takes the default (form 0: the tile).
// The tile is 12 waves tall: 156 KiB of LDS (140 with the learned guess, whose
// y results then share the side data's slots), one block per CU, 168 VGPRs
// (15-18 spilled; 39-77 in the G2D and learned forms) for 3 waves per SIMD.
// Against 8 waves (6 rows, 100 KiB, 192 VGPRs, 2 waves per SIMD) its halo rows
// are 1 in 6 instead of 1 in 4: at 1024^2 the whole run takes 602 instead of
// 696 ms (the reference's Newton setting), 667 instead of 706 ms (learned),
// 679 instead of 726 ms (G2D); 10 waves gave 626, 696 and 710 ms, 6 waves
// (3 waves per SIMD, 4 rows) 854 ms.
constexpr int TILE_W = 12;

// rows per strip of the strip form: enough strips to give every CU several waves
static int strip_rows(int m, int n)
{
    const long strips = (long)((n + 61) / 62);
    int R = 32;
    while (R > 4 && strips * ((m + R - 1) / R) < 4L * 2048) R /= 2;
    return R;
}

// Form 0 and 1: the tile; 2: the strip.  On MI355X the 12-wave tile is the
// faster at every size measured (four-quadrant, the reference's Newton
// setting, per step from t = 0: 1024^2 0.207 against 0.313 ms, 2048^2 0.698
// against 0.837 ms, 4096^2 2.467 against 2.88 ms).
static bool use_strip(int, int, int form) { return form == 2; }

int unsplit_max_rows() { return 65535 * 4; }   // grid.y of the strip form at R = 4

template <int NM, bool G2>
static hipError_t launch_unsplit(const UnsplitArgs &a, hipStream_t st)
{
    if (use_strip(a.m, a.n, a.form)) {
        const int R = strip_rows(a.m, a.n);
        const dim3 grid((unsigned)((a.n + 61) / 62), (unsigned)((a.m + R - 1) / R));
        if (G2) hipLaunchKernelGGL(unsplit_strip_g2d_kernel<NM>, grid, dim3(64), 0, st, a, R);
        else hipLaunchKernelGGL(unsplit_strip_kernel<NM>, grid, dim3(64), 0, st, a, R);
    } else {
        const dim3 grid((unsigned)((a.n + 61) / 62), (unsigned)((a.m + TILE_W - 3) / (TILE_W - 2)));
        if (G2) hipLaunchKernelGGL((unsplit_step_g2d_kernel<TILE_W, 1, NM>), grid, dim3(64 * TILE_W), 0, st, a);
        else hipLaunchKernelGGL((unsplit_step_kernel<TILE_W, 1, NM>), grid, dim3(64 * TILE_W), 0, st, a);
    }
    return hipGetLastError();
}

template <bool G2>
static hipError_t launch_unsplit_nm(const UnsplitArgs &a, hipStream_t st)
{
    if (a.mlp) return launch_unsplit<NM_LEARNED, G2>(a, st);
    return a.conv ? launch_unsplit<NM_CONV, G2>(a, st) : launch_unsplit<NM_REF, G2>(a, st);
}

hipError_t launch_unsplit_step(const UnsplitArgs &a, hipStream_t st)
{
    return a.g2d ? launch_unsplit_nm<true>(a, st) : launch_unsplit_nm<false>(a, st);
}

const char *unsplit_kernel_name(int m, int n, int nm, int g2d, int form)
{
    static const char *names[2][3][2] = {
        {{"unsplit_step_kernel<12, 1, 0>", "unsplit_strip_kernel<0>"},
         {"unsplit_step_kernel<12, 1, 1>", "unsplit_strip_kernel<1>"},
         {"unsplit_step_kernel<12, 1, 2>", "unsplit_strip_kernel<2>"}},
        {{"unsplit_step_g2d_kernel<12, 1, 0>", "unsplit_strip_g2d_kernel<0>"},
         {"unsplit_step_g2d_kernel<12, 1, 1>", "unsplit_strip_g2d_kernel<1>"},
         {"unsplit_step_g2d_kernel<12, 1, 2>", "unsplit_strip_g2d_kernel<2>"}}};
    if (nm < 0 || nm > 2) return "";
    return names[g2d ? 1 : 0][nm][use_strip(m, n, form) ? 1 : 0];
}

}  // namespace hgrp
