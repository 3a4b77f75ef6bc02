// sweep2d_body.hpp -- the bodies of the directional sweeps of the
// dimension-split 2D GRP step, shared by the one-grid kernels (sweep2d.hip)
// and the kernels that advance a batch of grids (grp2d_batch.hip).
//
// Semantics follow src/finite_volume/grp_solver_2D_split_EUL_source.c:182-303,
// src/inter_process/bound_cond_slope_limiter_{x,y}.c:27-129,
// src/inter_process/slope_limiter{,_2D_x}.c and src/flux_calc/flux_generator_{x,y}.c:31-165.
#pragma once
#include <hip/hip_runtime.h>

#include "grp_device.hpp"
#include "sweep2d.hpp"

namespace hgrp {

// The argument block of a one-grid sweep kernel, read again from
// kernel-argument memory (see reload_args).
struct KernargArgs {
    __device__ SweepArgs get() const { return reload_args<SweepArgs>(); }
};

__device__ __forceinline__ bool sweep_halted(const SweepArgs &a)
{
    if (a.prev == nullptr) return false;
    const SweepStatus pv = *a.prev;
    if (pv.fail_key != 0ull || pv.flags != 0u) {
        if (threadIdx.x == 0) atomicOr(&a.status->flags, 2u);
        return true;
    }
    return false;
}

// One grid line along the sweep direction.  DIR 0: x-line i (cells j = 0..m-1,
// stride ld); DIR 1: y-line j (cells i = 0..n-1, stride 1).  Index -1 and cnt
// are the ghost cells.
template <int DIR>
struct Line {
    const SweepArgs &a;
    size_t base, stride;
    int cnt, line;
    static constexpr int NV = DIR == 0 ? 1 : 2;   // normal velocity: u in x, v in y

    __device__ Line(const SweepArgs &args, int ln) : a(args), line(ln)
    {
        if (DIR == 0) { base = (size_t)ln; stride = (size_t)args.ld; cnt = args.m; }
        else          { base = (size_t)ln * (size_t)args.ld; stride = 1; cnt = args.n; }
    }
    __device__ size_t at(int k) const { return base + (size_t)k * stride; }

    __device__ void cell(int k, double w[4]) const
    {
        const size_t o = at(k);
        const double *f = a.st + o;
        w[0] = f[0]; w[1] = f[a.plane]; w[2] = f[2 * a.plane]; w[3] = f[3 * a.plane];
    }
    // value of cell k, ghost values at k == -1 / k == cnt (bound_cond_slope_limiter_x.c:34-82)
    __device__ void value(int k, double w[4]) const
    {
        if (k >= 0 && k < cnt) { cell(k, w); return; }
        const bool lo = k < 0;
        switch (a.bc) {
        case BC_INITIAL: {
            const double *b = (lo ? a.bnd_lo : a.bnd_hi) + (size_t)line * 4;
            w[0] = b[0]; w[1] = b[1]; w[2] = b[2]; w[3] = b[3];
            return;
        }
        case BC_PERIODIC:
            cell(lo ? cnt - 1 : 0, w);
            return;
        case BC_REFLECT:
            cell(lo ? 0 : cnt - 1, w);
            w[NV] = -w[NV];
            return;
        case BC_REFLECT_FREE:
            cell(lo ? 0 : cnt - 1, w);
            if (lo) w[NV] = -w[NV];
            return;
        default:   // BC_FREE
            cell(lo ? 0 : cnt - 1, w);
            return;
        }
    }
    // previous slopes of cell k (read by minmod3 only)
    __device__ void prev_slopes(int k, double sp[4]) const
    {
        const size_t o = at(k);
#pragma unroll
        for (int q = 0; q < 4; ++q) sp[q] = a.limiter3 ? a.s[o + q * a.plane] : 0.0;
    }
    // limited slope from the neighbours and the previous slope sp (slope_limiter_2D_x.c:177-211)
    __device__ void limited_sp(const double vm[4], const double v0[4], const double vp[4], const double sp[4],
                               double s[4]) const
    {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double s_L = div_by(v0[q] - vm[q], a.div_h);
            const double s_R = div_by(vp[q] - v0[q], a.div_h);
            s[q] = a.limiter3 ? minmod3(a.alpha * s_L, a.alpha * s_R, sp[q]) : minmod2(s_L, s_R);
        }
    }
    __device__ void limited(int k, const double vm[4], const double v0[4], const double vp[4], double s[4]) const
    {
        double sp[4];
        prev_slopes(k, sp);
        limited_sp(vm, v0, vp, sp, s);
    }
    __device__ void limited_at(int k, double s[4]) const
    {
        double vm[4], v0[4], vp[4];
        value(k - 1, vm); value(k, v0); value(k + 1, vp);
        limited(k, vm, v0, vp, s);
    }
    // ghost slopes (bound_cond_slope_limiter_x.c:94-109); zero where the reference leaves them unset
    __device__ void ghost_slope(bool lo, double s[4]) const
    {
        s[0] = s[1] = s[2] = s[3] = 0.0;
        double t[4];
        switch (a.bc) {
        case BC_REFLECT:
            limited_at(lo ? 0 : cnt - 1, t);
            s[NV] = t[NV];
            break;
        case BC_PERIODIC:
            limited_at(lo ? cnt - 1 : 0, s);
            break;
        case BC_REFLECT_FREE:
            if (lo) { limited_at(0, t); s[NV] = t[NV]; }
            break;
        default:
            break;
        }
    }
    // conservative update of cell k and slope refresh (grp_solver_2D_split_EUL_source.c:236-254)
    __device__ void update(int k, const double w[4], const double Fl[4], const double Fr[4],
                           const double Il[4], const double Ir[4]) const
    {
        update_e(k, w, a.st[at(k) + 4 * a.plane], Fl, Fr, Il, Ir);
    }
    // the same with E of the cell already loaded
    __device__ void update_e(int k, const double w[4], const double e, const double Fl[4], const double Fr[4],
                             const double Il[4], const double Ir[4]) const
    {
        const size_t o = at(k);
        const double nu = a.nu;
        const double mom_x = w[0] * w[1] - nu * (Fr[1] - Fl[1]);
        const double mom_y = w[0] * w[2] - nu * (Fr[2] - Fl[2]);
        const double ene = w[0] * e - nu * (Fr[3] - Fl[3]);
        const double rho = w[0] + -nu * (Fr[0] - Fl[0]);
        const double u = mom_x / rho;
        const double v = mom_y / rho;
        const double p = (ene - 0.5 * mom_x * u - 0.5 * mom_y * v) * a.gas.gm1();
        double *f = a.st_o + o;
        const size_t P = a.plane;
        f[0] = rho; f[P] = u; f[2 * P] = v; f[4 * P] = ene / rho; f[3 * P] = p;
        if (p < a.eps || rho < a.eps) atomicOr(&a.status->flags, 1u);
#pragma unroll
        for (int q = 0; q < 4; ++q) a.s_o[o + q * P] = div_by(Ir[q] - Il[q], a.div_h);
    }
};

template <int DIR, int NM>
__device__ __forceinline__ void solve_interface(const SweepArgs &a, const Side &SL, const Side &SR,
                                                unsigned long long loop_index, double F[4], double I[4],
                                                double p_guess = -1.0)
{
    int code = ifvar_bad(SL, SR, a.eps) ? 1 : 0;
    if (grp_2d_flux<DIR, NM != NM_REF>(SL, SR, a.gas, a.tau, a.eps, F, I, a.tol, p_guess)) code = 2;
    if (code) atomicMax(&a.status->fail_key, ((loop_index << 2) | (unsigned long long)code) + 1ull);
}

// Learned starting pressure for the interface (SL | SR) normal to DIR, for
// every lane of the wave (uniform call); -1 where Newton will not run (the
// Q1D acoustic shortcut dist < eps, the solver's equal-state exit) or the lane
// has no interface.
template <int DIR>
__device__ __forceinline__ double interface_guess(const SweepArgs &a, bool active, const Side &SL, const Side &SR,
                                                  uint32_t *lds)
{
    const double uL = DIR == 0 ? SL.u : SL.v, uR = DIR == 0 ? SR.u : SR.v;
    const double dist = sqrt((SL.rho - SR.rho) * (SL.rho - SR.rho) + (uL - uR) * (uL - uR)
                             + (SL.p - SR.p) * (SL.p - SR.p));
    // Newton runs only past the acoustic shortcut and the equal-state exit of the exact solver
    const bool want = active && !(dist < a.eps) && !(fabs(uL - uR) < a.tol && fabs(SL.p - SR.p) < a.tol);
    return learned_guess_wave(a.mlp, want, SL.rho, uL, SL.p, SR.rho, uR, SR.p, a.gas.g, a.gas.g, lds);
}

// Line-parallel sweep along DIR: a block of W waves covers 64 consecutive
// grid lines of the sweep direction (lane: line; DIR 0: x-line i, DIR 1:
// y-line j) and the cells c0-1 .. c0+W-2 along them (wave w: cell c = c0-1+w).
// Neighbours along the sweep live in the next wave, so values, slopes and then
// fluxes are exchanged through LDS (one array each: three barriers per block).
// Wave w solves interface c+1 (between cells c and c+1); waves 1..W-1 update
// cells c0 .. c0+W-2.  The 64 lanes of a wave lie across the sweep, so an
// interface front (a shock or contact crossing the lines) is met by the whole
// wave together and the GRP branches stay coherent.  Used for DIR 0, where the
// lanes read consecutive doubles.  (For DIR 1 each lane would read its own row,
// stride ld: on MI355X that form ran as fast as sweep_y_kernel at 1024^2 and
// 3x slower at 4096^2, where the rows no longer stay in cache.)
// One chunk of a line-parallel sweep: the block's waves own cells
// c0 - 1 + w (w = 0 .. W-1) of the 64 lines of group bx.  Wave w solves the
// face right of its cell; waves 1 .. W-1 update their cells with the face left
// of them, and wave 0 too when `carried` (with the face the previous chunk's
// last wave left in xc[(k - 1) & 1]); with `carry` the last wave leaves its
// face in xc[k & 1].  Returns false when the sweep is halted (only checked
// with `first`).
// lane, w: the thread's lane and wave (threadIdx.x & 63, threadIdx.x >> 6).
// The global loads of a chunk for the thread's own cell c = c0 - 1 + w: its
// values, previous slopes and E.
struct ChunkIn {
    double v[4], sp[4], e_c;
};

template <int DIR, int W>
__device__ __forceinline__ void chunk_loads(const SweepArgs &a, int lane, int w, int bx, int c0, ChunkIn &in)
{
    const int line = bx * 64 + lane;
    const bool col = line < (DIR == 0 ? a.n : a.m);
    const Line<DIR> L(a, col ? line : 0);
    const int m = L.cnt;
    const int c = c0 - 1 + w;
    const bool own = col && c >= 0 && c < m;
    in.v[0] = 1.0; in.v[1] = 0.0; in.v[2] = 0.0; in.v[3] = 1.0;
    in.sp[0] = in.sp[1] = in.sp[2] = in.sp[3] = 0.0;
    in.e_c = 0.0;
    if (col && c <= m) L.value(c, in.v);
    if (own) {
        L.prev_slopes(c, in.sp);
        in.e_c = a.st[L.at(c) + 4 * a.plane];
    }
}

template <int DIR, int W, int NM, class Src>
__device__ __forceinline__ bool line_chunk(const SweepArgs &a, const Src &src, int lane, int w, int bx, int c0, int k,
                                          bool first, bool carried, bool carry, double (*xa)[4][64],
                                          double (*xb)[4][64], double (*xf)[4][64], double (*xi)[4][64],
                                          double (*xe)[64], double (*xc)[8][64])
{
    const int line = bx * 64 + lane;
    const bool col = line < (DIR == 0 ? a.n : a.m);
    const Line<DIR> L(a, col ? line : 0);
    const int m = L.cnt;                          // cells along the sweep
    const int c = c0 - 1 + w;                     // owned cell; -1 and m are ghosts
    const double hh = a.hh;

    // every global load of the chunk up front, before the first LDS store:
    // the owned cell's values, previous slopes and E, then the end waves'
    // extra cells
    const bool own = col && c >= 0 && c < m;
    double v[4], vl[4], vr[4], sp[4], e_c;
    {
        ChunkIn in;
        chunk_loads<DIR, W>(a, lane, w, bx, c0, in);
        e_c = in.e_c;
#pragma unroll
        for (int q = 0; q < 4; ++q) { v[q] = in.v[q]; sp[q] = in.sp[q]; }
    }
    if (col && w == 0 && c >= 0) L.value(c - 1, vl);
    // the last wave also limits cell c + 1 (the right side of its interface): its loads go out now too
    double vrr[4], sp1[4];
    if (col && w == W - 1 && c + 1 <= m) L.value(c + 1, vr);
    if (col && w == W - 1 && c + 1 < m) {
        L.value(c + 2, vrr);
        L.prev_slopes(c + 1, sp1);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) xa[w][q][lane] = v[q];
    xe[w][lane] = e_c;
    const bool next_own = col && w == W - 1 && c + 1 < m;
    // the predecessor's status is read while the loads above are in flight
    // (they stay in bounds either way); the whole block returns together
    if (first && sweep_halted(a)) return false;
    __syncthreads();
    if (w > 0)
#pragma unroll
        for (int q = 0; q < 4; ++q) vl[q] = xa[w - 1][q][lane];
    if (w < W - 1)
#pragma unroll
        for (int q = 0; q < 4; ++q) vr[q] = xa[w + 1][q][lane];

    double s[4] = {0.0, 0.0, 0.0, 0.0}, sr[4] = {0.0, 0.0, 0.0, 0.0};
    if (col) {
        if (own) L.limited_sp(vl, v, vr, sp, s);
        else if (c == -1) L.ghost_slope(true, s);
        else if (c == m) L.ghost_slope(false, s);
        if (w == W - 1) {
            if (next_own) {
                L.limited_sp(v, vr, vrr, sp1, sr);
            } else if (c + 1 == m) {
                L.ghost_slope(false, sr);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) xb[w][q][lane] = s[q];
    __syncthreads();
    if (w < W - 1)
#pragma unroll
        for (int q = 0; q < 4; ++q) sr[q] = xb[w + 1][q][lane];

    double F[4] = {0.0, 0.0, 0.0, 0.0}, I[4] = {0.0, 0.0, 0.0, 0.0};
    const int f = c + 1;
    double pg = -1.0;
    if constexpr (NM == NM_LEARNED) {
        // xf[w] is not written before the solve: it holds this wave's network inputs
        const Side SL = recon(v, s, hh, true);
        const Side SR = recon(vr, sr, hh, false);
        pg = interface_guess<DIR>(src.get(), col && f <= m, SL, SR, (uint32_t *)&xf[w][0][0]);
    }
    if (col && f <= m) {
        // the argument block is read again here and after the solve, so that
        // none of it is held across the solve
        const Side SL = recon(v, s, hh, true);
        const Side SR = recon(vr, sr, hh, false);
        solve_interface<DIR, NM>(src.get(), SL, SR,
                                 (unsigned long long)line * (unsigned long long)(m + 1) + (unsigned long long)f, F, I, pg);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) { xf[w][q][lane] = F[q]; xi[w][q][lane] = I[q]; }
    if (carry && w == W - 1)                      // the next chunk's first face
#pragma unroll
        for (int q = 0; q < 4; ++q) { xc[k & 1][q][lane] = F[q]; xc[k & 1][4 + q][lane] = I[q]; }
    __syncthreads();
    if (own && (w >= 1 || carried)) {
        // the cell's values and E come back from LDS, so no register holds them across the solve
        const SweepArgs ar = src.get();
        const Line<DIR> Lr(ar, line);
        double Fl[4], Il[4], vc[4];
        if (w >= 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) { Fl[q] = xf[w - 1][q][lane]; Il[q] = xi[w - 1][q][lane]; }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) { Fl[q] = xc[(k - 1) & 1][q][lane]; Il[q] = xc[(k - 1) & 1][4 + q][lane]; }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) vc[q] = xa[w][q][lane];
        Lr.update_e(c, vc, xe[w][lane], Fl, F, Il, I);
    }
    return true;
}

// Line-parallel sweep of one block: (bx, by) its group of 64 lines and its
// run of W - 1 cells (by (W - 1) onwards, wave 0 holding the halo cell); src
// gives the argument block again after the solve.
template <int DIR, int W, int NM, class Src>
__device__ __forceinline__ void line_sweep(const SweepArgs &a, const Src &src, int bx, int by, double (*xa)[4][64],
                                           double (*xb)[4][64], double (*xf)[4][64], double (*xi)[4][64],
                                           double (*xe)[64])
{
    (void)line_chunk<DIR, W, NM>(a, src, threadIdx.x & 63, threadIdx.x >> 6, bx, by * (W - 1), 0, true, false, false,
                                 xa, xb, xf, xi, xe, nullptr);
}

// The same walking K chunks of W cells along the block's lines (by (K W - 1)
// onwards): the first chunk's wave 0 is the halo cell as above, and in every
// later chunk wave 0 updates its cell too, with the face that the previous
// chunk's last wave solved (xc, double-buffered); a block then solves K W faces
// for K W - 1 cells instead of W for W - 1.  Every LDS slot a chunk writes before
// its first barrier is read by the previous chunk only before that chunk's last
// barrier, so the chunks need no barrier between them.  The argument block is
// read again for each chunk, and the lane and wave indices pass through an
// empty asm each time, so that neither the arguments nor the LDS addresses
// derived from them are held across the chunks' solves.
template <int DIR, int W, int NM, int K, class Src>
__device__ __forceinline__ void line_sweep_chunks(const Src &src, int bx, int by, double (*xa)[4][64],
                                                  double (*xb)[4][64], double (*xf)[4][64], double (*xi)[4][64],
                                                  double (*xe)[64], double (*xc)[8][64])
{
    const int s0 = by * (K * W - 1);
    for (int k = 0; k < K; ++k) {
        const SweepArgs a = src.get();
        const int c0 = s0 + k * W;
        if (c0 - 1 >= (DIR == 0 ? a.m : a.n)) break;   // uniform: the chunk lies past the last cell
        int lane = threadIdx.x & 63;
        int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        __asm__ volatile("" : "+v"(lane));
        __asm__ volatile("" : "+s"(w));
        if (!line_chunk<DIR, W, NM>(a, src, lane, w, bx, c0, k, k == 0, k > 0, k + 1 < K, xa, xb, xf, xi, xe, xc))
            return;
    }
}

// Y sweep: a wave holds 64/G y-lines (rows j) of G lanes each; lane k of a
// row owns cell c0-1+k, computes interface c0+k and updates cell c0-1+k for
// k >= 1 (G-1 cells per row).  With G = 64 a wave is one stretch of one line;
// smaller G stacks rows, so that a front lying along y (constant i) meets
// fewer waves, each with more lanes on it.
// wave: the wave's index in the grid's sweep (a group of 64/G rows, one run
// of G - 1 cells); st, mlp: this wave's LDS slots; src gives the argument
// block again after the solve.
template <int NM, int G, class Src>
__device__ __forceinline__ void y_sweep(const SweepArgs &a, const Src &src, long wave, int chunks,
                                        double (*st)[64], uint32_t *mlp)
{
    static_assert(G == 64 || G == 32 || G == 16 || G == 8, "lanes per row");
    constexpr int ROWS = 64 / G;
    constexpr bool LEARNED = NM == NM_LEARNED;
    const int lane = threadIdx.x & 63, k = lane % G;
    const int jg = (int)(wave / chunks);
    if (jg * ROWS >= a.m) return;                      // uniform per wave
    const int j = jg * ROWS + lane / G;
    const bool rowv = j < a.m;
    const int c0 = (int)(wave % chunks) * (G - 1);
    const int n = a.n;
    const int c = c0 - 1 + k;
    const Line<1> L(a, rowv ? j : 0);
    const double hh = a.hh;

    const bool own = rowv && c >= 0 && c < n;
    // every global load of the wave up front, so that it pays one memory
    // latency: the owned cell (values, previous slopes, E); the first lane of a
    // row also the cell below it; the last lane the two cells above it and the
    // previous slopes of the first (it limits that cell, the right side of its
    // interface)
    double w[4] = {1.0, 0.0, 0.0, 1.0}, sp[4] = {0.0, 0.0, 0.0, 0.0}, e_c = 0.0;
    if (rowv && c <= n) L.value(c, w);
    if (own) {
        L.prev_slopes(c, sp);
        e_c = a.st[L.at(c) + 4 * a.plane];
    }
    const bool first = rowv && k == 0 && c >= 0, last = rowv && k == G - 1 && c + 1 <= n;
    const bool last_own = rowv && k == G - 1 && c + 1 < n;
    double wl_e[4], wr_e[4], wrr[4], sp1[4];
    if (first) L.value(c - 1, wl_e);
    if (last) L.value(c + 1, wr_e);
    if (last_own) {
        L.value(c + 2, wrr);
        L.prev_slopes(c + 1, sp1);
    }
    // the predecessor's status is read while the loads above are in flight
    if (sweep_halted(a)) return;
    // each lane reads back only its own slots: no barrier needed
#pragma unroll
    for (int q = 0; q < 4; ++q) st[q][lane] = w[q];
    st[4][lane] = e_c;
    // neighbours along the row from the next lanes; a row's end lanes take the loaded ones
    double wl[4], wr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { wl[q] = __shfl_up(w[q], 1); wr[q] = __shfl_down(w[q], 1); }
    if (first)
#pragma unroll
        for (int q = 0; q < 4; ++q) wl[q] = wl_e[q];
    if (last)
#pragma unroll
        for (int q = 0; q < 4; ++q) wr[q] = wr_e[q];

    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (own) L.limited_sp(wl, w, wr, sp, s);
    else if (rowv && c == -1) L.ghost_slope(true, s);
    else if (rowv && c == n) L.ghost_slope(false, s);

    double sr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) sr[q] = __shfl_down(s[q], 1);
    if (rowv && k == G - 1) {
        if (last_own) L.limited_sp(w, wr, wrr, sp1, sr);
        else if (c + 1 == n) L.ghost_slope(false, sr);
    }

    double F[4] = {0.0, 0.0, 0.0, 0.0}, I[4] = {0.0, 0.0, 0.0, 0.0};
    const int f = c + 1;
    const bool face = rowv && f <= n;
    double pg = -1.0;
    if constexpr (LEARNED) {
        const Side SL = recon(w, s, hh, true);
        const Side SR = recon(wr, sr, hh, false);
        pg = interface_guess<1>(src.get(), face, SL, SR, mlp);
    }
    if (face) {
        // the argument block is read again here and after the solve, so that
        // none of it is held across the solve
        const Side SL = recon(w, s, hh, true);
        const Side SR = recon(wr, sr, hh, false);
        solve_interface<1, NM>(src.get(), SL, SR,
                               (unsigned long long)j * (unsigned long long)(n + 1) + (unsigned long long)f, F, I, pg);
    }
    double Fl[4], Il[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { Fl[q] = __shfl_up(F[q], 1); Il[q] = __shfl_up(I[q], 1); }
    if (k >= 1 && own) {
        const SweepArgs ar = src.get();
        const Line<1> Lr(ar, j);
        double wc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) wc[q] = st[q][lane];
        Lr.update_e(c, wc, st[4][lane], Fl, F, Il, I);
    }
}

}  // namespace hgrp
