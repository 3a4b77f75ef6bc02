// grp2d_batch.hip -- a batch of independent 2D split GRP problems of uneven
// sizes advanced together: each sweep of every grid of the batch is one
// launch, each grid with its own frozen time step.  C ABI: hgrp2db_* in
// include/hydrocode_grp.h.
//
// Every grid runs GRP_solver_2D_split_EUL_source
// (src/finite_volume/grp_solver_2D_split_EUL_source.c:75-379) with the plot
// times {0, t_all}: tau = CFL h/S_max of its initial data (:183-191), clipped
// so that the run ends at t_all (:193-207); X(tau/2) Y(tau) X(tau/2) per step,
// the last step ending after its Y sweep (:299-304); the same stop and abort
// rules.  The schedule of every grid is the same sequence of sweep directions
// (X Y X X Y X ...), only its length differs, so all grids move through it
// together: sweep g of the batch is sweep g of every grid that has one, and
// the buffer sets all grids read and write swap in step.  Grids are ordered by
// their number of sweeps, longest first, so the grids still running are a
// prefix of that order and each launch covers a prefix of one tile table per
// kernel shape.  Each grid has its own ring of status slots: a failing
// interface or cell halts that grid's later sweeps on the device, and the
// host, reading the rings every STATUS_SLOTS sweeps, settles it with the
// reference's rule and takes it out of later launches.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/hydrocode_grp.h"
#include "grp_device.hpp"
#include "sweep2d.hpp"
#include "sweep2d_body.hpp"

namespace hgrp {

int sweep_variants(int dir);
hipError_t launch_cfl(const double *rho, const double *u, const double *v, const double *p, int m, int n,
                      int ld, double gamma, double hmin, unsigned long long *out, hipStream_t st);

// One grid of a batch: what differs between the grids.
struct BatchGrid {
    double *st[2];                 // state sets: rho u v p e, `plane` doubles apart
    double *sx[2], *sy[2];         // x- and y-slope sets: rho u v p
    const double *bnd[4];          // BC -1 ghosts from the initial data: x_lo, x_hi (n x 4), y_lo, y_hi (m x 4)
    size_t plane;
    int m, n, ld;                  // cells in x and in y, row pitch
    int last_step;                 // the step that takes tau_last
    int halted;                    // set by the host once the grid's run has ended on a failure
    int pad;
    double h_x, h_y, tau, tau_last;
    RecipDiv div_x, div_y;
    SweepStatus *ring;             // the grid's status slots
};

// What one launch of a batch sweep shares.
struct BatchSweep {
    const BatchGrid *grids;
    const int4 *tiles;             // per block: grid, then the kernel's block coordinates
    int dir, in, sin, limiter3;    // direction; state and slope sets read; minmod3 (a direction's later sweeps)
    int step, slot, prev_slot;     // the step's index, status slot of this sweep, of the one before (-1: none)
    int bc, conv;                  // boundary condition along the sweep, Newton rule
    double eps, alpha, tol;
    const MlpPacked *mlp;
    Gas gas;
};

// The one-grid argument block of grid gi for this sweep.  The time step and
// the update factors are formed as the one-grid loop forms them.
__device__ __forceinline__ SweepArgs batch_args(const BatchSweep &b, int gi)
{
    const BatchGrid &g = b.grids[gi];
    SweepArgs a = {};
    const bool x = b.dir == 0;
    a.st = g.st[b.in];
    a.st_o = g.st[1 - b.in];
    a.s = x ? g.sx[b.sin] : g.sy[b.sin];
    a.s_o = x ? g.sx[1 - b.sin] : g.sy[1 - b.sin];
    a.plane = g.plane;
    a.bnd_lo = g.bnd[x ? 0 : 2];
    a.bnd_hi = g.bnd[x ? 1 : 3];
    a.m = g.m; a.n = g.n; a.ld = g.ld;
    a.bc = b.bc;
    a.limiter3 = b.limiter3;
    a.conv = b.conv;
    const double tau = b.step == g.last_step ? g.tau_last : g.tau;
    if (x) {
        a.h = g.h_x;
        a.div_h = g.div_x;
        a.tau = tau * 0.5;
        a.nu = a.tau / g.h_x;
    } else {
        a.h = g.h_y;
        a.div_h = g.div_y;
        a.tau = tau;
        a.nu = tau / g.h_y;
    }
    a.hh = 0.5 * a.h;
    a.eps = b.eps; a.alpha = b.alpha; a.tol = b.tol;
    a.mlp = b.mlp;
    a.gas = b.gas;
    a.status = g.ring + b.slot;
    a.prev = b.prev_slot >= 0 ? g.ring + b.prev_slot : nullptr;
    return a;
}

// the argument block again after the solve: the launch's block from
// kernel-argument memory, the grid's entry from the table
struct BatchArgs {
    int gi;
    __device__ SweepArgs get() const { return batch_args(reload_args<BatchSweep>(), gi); }
};

template <int W, int MINW, int NM>
__global__ void __launch_bounds__(64 * W, MINW) batch_x_kernel(BatchSweep b)
{
    __shared__ double xa[W][4][64];
    __shared__ double xb[W][4][64];
    __shared__ double xf[W][4][64];
    __shared__ double xi[W][4][64];
    __shared__ double xe[W][64];
    const int4 t = b.tiles[blockIdx.x];
    if (b.grids[t.x].halted) return;                    // the whole block
    const SweepArgs a = batch_args(b, t.x);
    line_sweep<0, W, NM>(a, BatchArgs{t.x}, t.y, t.z, xa, xb, xf, xi, xe);
}

// the X sweep walking K chunks of W rows per block (line_sweep_chunks): the
// tile's row block is by (K W - 1) rows
template <int W, int MINW, int NM, int K>
__global__ void __launch_bounds__(64 * W, MINW) batch_x_chunks_kernel(BatchSweep b)
{
    __shared__ double xa[W][4][64];
    __shared__ double xb[W][4][64];
    __shared__ double xf[W][4][64];
    __shared__ double xi[W][4][64];
    __shared__ double xe[W][64];
    __shared__ double xc[2][8][64];                     // the last face of the previous chunk: F, I
    const int4 t = b.tiles[blockIdx.x];
    if (b.grids[t.x].halted) return;                    // the whole block
    line_sweep_chunks<0, W, NM, K>(BatchArgs{t.x}, t.y, t.z, xa, xb, xf, xi, xe, xc);
}

template <int MINW, int NM, int G>
__global__ void __launch_bounds__(256, MINW) batch_y_kernel(BatchSweep b)
{
    constexpr bool LEARNED = NM == NM_LEARNED;
    __shared__ uint32_t mlp_lds[LEARNED ? 4 : 1][LEARNED ? 64 * 8 : 1];
    __shared__ double stash[4][5][64];
    const int4 t = b.tiles[blockIdx.x];                 // grid, first wave of the block, chunks per row group
    if (b.grids[t.x].halted) return;
    const SweepArgs a = batch_args(b, t.x);
    const int wv = (threadIdx.x >> 6) & 3;
    y_sweep<NM, G>(a, BatchArgs{t.x}, (long)t.y + wv, t.z, stash[wv], mlp_lds[wv]);
}

// the one-grid launch variants (sweep2d.hip) X 0..4 and Y 0..4: X 0 = 8 rows
// at 4 waves/SIMD, 1 = 8 rows at the compiler's budget, 2 = 16 rows, 3 and 4 =
// 8 rows walking 4 and 8 chunks (31 and 63 rows per block); Y 0..4 = 16, 32,
// 64, 8 lanes per row at 4 waves/SIMD, 64 at the compiler's budget
static int x_cells_per_block(int v) { return v == 2 ? 15 : v == 3 ? 31 : v == 4 ? 63 : 7; }
static int y_lanes_b(int v) { return v == 1 ? 32 : (v == 2 || v == 4) ? 64 : v == 3 ? 8 : 16; }

template <int NM>
static void launch_bx(const BatchSweep &b, int v, int blocks, hipStream_t st)
{
    switch (v) {
    case 1: hipLaunchKernelGGL((batch_x_kernel<8, 1, NM>), dim3(blocks), dim3(512), 0, st, b); break;
    case 2: hipLaunchKernelGGL((batch_x_kernel<16, 4, NM>), dim3(blocks), dim3(1024), 0, st, b); break;
    case 3: hipLaunchKernelGGL((batch_x_chunks_kernel<8, 4, NM, 4>), dim3(blocks), dim3(512), 0, st, b); break;
    case 4: hipLaunchKernelGGL((batch_x_chunks_kernel<8, 4, NM, 8>), dim3(blocks), dim3(512), 0, st, b); break;
    default: hipLaunchKernelGGL((batch_x_kernel<8, 4, NM>), dim3(blocks), dim3(512), 0, st, b); break;
    }
}

template <int NM>
static void launch_by(const BatchSweep &b, int v, int blocks, hipStream_t st)
{
    switch (v) {
    case 1: hipLaunchKernelGGL((batch_y_kernel<4, NM, 32>), dim3(blocks), dim3(256), 0, st, b); break;
    case 2: hipLaunchKernelGGL((batch_y_kernel<4, NM, 64>), dim3(blocks), dim3(256), 0, st, b); break;
    case 3: hipLaunchKernelGGL((batch_y_kernel<4, NM, 8>), dim3(blocks), dim3(256), 0, st, b); break;
    case 4: hipLaunchKernelGGL((batch_y_kernel<1, NM, 64>), dim3(blocks), dim3(256), 0, st, b); break;
    default: hipLaunchKernelGGL((batch_y_kernel<4, NM, 16>), dim3(blocks), dim3(256), 0, st, b); break;
    }
}

static hipError_t launch_batch_sweep(const BatchSweep &b, int nm, int v, int blocks, hipStream_t st)
{
    if (blocks <= 0) return hipSuccess;
    if (b.dir == 0) {
        if (nm == NM_LEARNED) launch_bx<NM_LEARNED>(b, v, blocks, st);
        else if (nm == NM_CONV) launch_bx<NM_CONV>(b, v, blocks, st);
        else launch_bx<NM_REF>(b, v, blocks, st);
    } else {
        if (nm == NM_LEARNED) launch_by<NM_LEARNED>(b, v, blocks, st);
        else if (nm == NM_CONV) launch_by<NM_CONV>(b, v, blocks, st);
        else launch_by<NM_REF>(b, v, blocks, st);
    }
    return hipGetLastError();
}

}  // namespace hgrp

using namespace hgrp;

#define BCHECK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "hydrogrp: %s failed: %s (%s:%d)\n", #x,  \
                    hipGetErrorString(e_), __FILE__, __LINE__);       \
            return -(int)e_ - 1;                                      \
        }                                                             \
    } while (0)

namespace {

constexpr int RING = 96;             // sweeps between host checks
constexpr int ARG = -1000;
constexpr int NX_VAR = 5, NY_VAR = 5;

// One grid's run as the reference's loop takes it when nothing fails: its
// sweeps, its two time steps, and how it ends.
struct Sched {
    int sweeps = 0;                  // sweeps the run takes
    int last_step = -1;              // the step that takes tau_last
    double tau = 0.0, tau_last = 0.0;
    std::vector<double> t;           // time_c before step s (t[s]), s = 0 .. steps
    int k = 1, status = 0;           // config[5] and the status at the end
    double time_c = 0.0, tau_end = 0.0;
};

bool valid_bc(int b) { return b == -1 || b == -2 || b == -4 || b == -7 || b == -24; }

// The control flow of hgrp2d_solve (GRP_solver_2D_split_EUL_source:147-341)
// for time_plot = {0, t_all} without failures; h_S_max of the initial data.
Sched schedule(const double *cfg, double h_S_max)
{
    Sched S;
    const double t_all = cfg[1], eps = cfg[4], CFL = cfg[7];
    const int N = cfg[5] >= 1.0 && cfg[5] < 2147483648.0 ? (int)cfg[5] : 2147483647;
    double tau = cfg[16], time_c = 0.0;
    int k = 1, DS = 1, status = 0, s = 0;
    bool stop_t = false;
    S.t.push_back(0.0);
    for (; k <= N; DS ? k : ++k) {
        if (DS) {
            if (isfinite(t_all) || !isfinite(cfg[16]) || cfg[16] <= 0.0) {
                tau = CFL * h_S_max;
                if (tau < eps) {
                    stop_t = true;
                } else if ((time_c + tau) > (t_all - eps)) {
                    tau = t_all - time_c;
                } else if (!isfinite(tau)) {
                    status = 1;
                    break;
                }
            }
            if (s == 0) S.tau = tau;
            else if (tau != S.tau && S.last_step < 0) S.last_step = s;
            if (S.last_step == s) S.tau_last = tau;
        }
        ++S.sweeps;                                       // X(tau/2)
        if (stop_t) break;
        if (DS) {
            ++S.sweeps;                                   // Y(tau)
            time_c += tau;
            ++s;
            S.t.push_back(time_c);
        }
        if (time_c > (t_all - eps) || !isfinite(time_c)) break;
        DS = DS ? 0 : 1;
    }
    if (stop_t && status == 0) status = 2;
    if (S.last_step < 0) { S.last_step = -1; S.tau_last = S.tau; }
    S.k = k;
    S.status = status;
    S.time_c = time_c;
    S.tau_end = tau;
    return S;
}

}  // namespace

struct hgrp2db {
    int count;
    std::vector<int> m, n, ld;
    std::vector<double> hx, hy;
    std::vector<size_t> base;        // first double of each grid in `block`
    double cfg[HGRP_N_CONF];
    hipStream_t stream;
    double *block = nullptr;
    size_t block_doubles = 0;
    BatchGrid *grids_d = nullptr;
    std::vector<BatchGrid> grids;
    SweepStatus *status = nullptr;   // count x RING slots
    unsigned long long *cfl_bits = nullptr;
    std::vector<double> h_S_max;
    std::vector<Sched> sched;
    std::vector<int> order;          // grids by number of sweeps, longest first
    int4 *tiles_d[2][5] = {};
    std::vector<int> prefix[2][5];   // tiles of the first r grids of `order`
    int xv = 0, yv = 0;
    double newton_tol = 0.0;
    int newton_rule = HGRP_NEWTON_REFERENCE;
    MlpPacked *mlp = nullptr;
    bool use_mlp = false;
    bool uploaded = false;
    // results
    std::vector<int> res_k, res_status, res_cs, res_steps;
    std::vector<double> res_time;
    hipEvent_t ev[2];
};

static size_t pitch_of(int n) { return (size_t)((n + 31) / 32) * 32; }

extern "C" int hgrp2db_destroy(hgrp2db *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->block) (void)hipFree(c->block);
    if (c->grids_d) (void)hipFree(c->grids_d);
    if (c->status) (void)hipFree(c->status);
    for (auto &d : c->tiles_d)
        for (auto *t : d)
            if (t) (void)hipFree(t);
    if (c->mlp) (void)hipFree(c->mlp);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

extern "C" int hgrp2db_create(hgrp2db **out, int count, const int *m, const int *n, const double *h_x,
                              const double *h_y, const double *config)
{
    if (!out || count < 1 || !m || !n || !h_x || !h_y || !config) return ARG - 1;
    *out = nullptr;
    for (int i = 0; i < count; ++i)
        if (m[i] < 1 || n[i] < 1 || !(h_x[i] > 0.0) || !(h_y[i] > 0.0)) return ARG - 2;
    if (!valid_bc((int)config[17]) || !valid_bc((int)config[18])) return ARG - 4;
    // the schedule is formed up front: the run must end by time or by step count
    if (!isfinite(config[1]) && !(config[5] >= 1.0 && config[5] < 2147483648.0)) return ARG - 3;
    hgrp2db *c = new hgrp2db();
    c->count = count;
    memcpy(c->cfg, config, sizeof c->cfg);
    c->m.assign(m, m + count);
    c->n.assign(n, n + count);
    c->hx.assign(h_x, h_x + count);
    c->hy.assign(h_y, h_y + count);
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        c->ld.push_back((int)pitch_of(n[i]));
        c->base.push_back(total);
        total += (size_t)26 * m[i] * c->ld[i] + (size_t)8 * (m[i] + n[i]);
    }
    c->block_doubles = total;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&c->block, total * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&c->grids_d, (size_t)count * sizeof(BatchGrid));
    if (e == hipSuccess)
        e = hipMalloc((void **)&c->status, (size_t)count * RING * sizeof(SweepStatus) + (size_t)count * 8);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[1]);
    if (e == hipSuccess) e = hipMemsetAsync(c->block, 0, total * sizeof(double), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "hydrogrp: batch of %d grids, %.1f MiB: %s\n", count, total * 8.0 / 1048576.0,
                hipGetErrorString(e));
        hgrp2db_destroy(c);
        return -(int)e - 1;
    }
    c->cfl_bits = (unsigned long long *)(c->status + (size_t)count * RING);
    c->grids.resize(count);
    for (int i = 0; i < count; ++i) {
        BatchGrid &g = c->grids[i];
        const size_t plane = (size_t)m[i] * c->ld[i];
        double *q = c->block + c->base[i];
        g.st[0] = q; g.st[1] = q + 5 * plane;
        g.sx[0] = q + 10 * plane; g.sx[1] = q + 14 * plane;
        g.sy[0] = q + 18 * plane; g.sy[1] = q + 22 * plane;
        double *b = q + 26 * plane;
        g.bnd[0] = b; g.bnd[1] = b + (size_t)n[i] * 4; g.bnd[2] = b + (size_t)n[i] * 8;
        g.bnd[3] = b + (size_t)n[i] * 8 + (size_t)m[i] * 4;
        g.plane = plane;
        g.m = m[i]; g.n = n[i]; g.ld = c->ld[i];
        g.h_x = h_x[i]; g.h_y = h_y[i];
        g.div_x = recip_div(h_x[i]);
        g.div_y = recip_div(h_y[i]);
        g.ring = c->status + (size_t)i * RING;
    }
    *out = c;
    return 0;
}

extern "C" int hgrp2db_set_newton(hgrp2db *c, double tol, int rule, const float *params)
{
    if (!c || !(tol == tol) || (rule != HGRP_NEWTON_REFERENCE && rule != HGRP_NEWTON_CONVERGE)) return ARG - 1;
    if (params && rule != HGRP_NEWTON_CONVERGE) return ARG - 2;
    BCHECK(hipStreamSynchronize(c->stream));
    c->newton_tol = tol;
    c->newton_rule = rule;
    c->use_mlp = false;
    if (params) {
        MlpPacked h;
        mlp_pack_host(params, &h);
        if (!c->mlp) BCHECK(hipMalloc((void **)&c->mlp, sizeof(MlpPacked)));
        BCHECK(hipMemcpy(c->mlp, &h, sizeof h, hipMemcpyHostToDevice));
        c->use_mlp = true;
    }
    return 0;
}

extern "C" int hgrp2db_set_sweeps(hgrp2db *c, int xv, int yv)
{
    if (!c || xv < 0 || xv >= NX_VAR || yv < 0 || yv >= NY_VAR) return ARG - 1;
    c->xv = xv;
    c->yv = yv;
    return 0;
}

// the tile table of kernel shape (dir, v) over the grids in `order`
static int build_tiles(hgrp2db *c, int dir, int v)
{
    if (c->tiles_d[dir][v]) return 0;
    std::vector<int4> t;
    std::vector<int> &pre = c->prefix[dir][v];
    pre.assign(1, 0);
    for (int r = 0; r < c->count; ++r) {
        const int i = c->order[r], m = c->m[i], n = c->n[i];
        if (dir == 0) {
            const int rows = x_cells_per_block(v);
            for (int by = 0; by < (m + rows - 1) / rows; ++by)
                for (int bx = 0; bx < (n + 63) / 64; ++bx) t.push_back(make_int4(i, bx, by, 0));
        } else {
            const int G = y_lanes_b(v), rows = 64 / G;
            const int chunks = (n + G - 2) / (G - 1);
            const long waves = (long)((m + rows - 1) / rows) * chunks;
            for (long w = 0; w < waves; w += 4) t.push_back(make_int4(i, (int)w, chunks, 0));
        }
        pre.push_back((int)t.size());
    }
    BCHECK(hipMalloc((void **)&c->tiles_d[dir][v], std::max<size_t>(t.size(), 1) * sizeof(int4)));
    BCHECK(hipMemcpy(c->tiles_d[dir][v], t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int hgrp2db_upload(hgrp2db *c, const double *rho, const double *u, const double *v, const double *p)
{
    if (!c || !rho || !u || !v || !p) return ARG - 1;
    const double gamma = c->cfg[6];
    BCHECK(hipMemsetAsync(c->block, 0, c->block_doubles * sizeof(double), c->stream));
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    std::vector<unsigned long long> infs((size_t)c->count, inf_bits);
    BCHECK(hipMemcpyAsync(c->cfl_bits, infs.data(), infs.size() * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<std::vector<double>> keep;          // host buffers alive until the copies are done
    size_t off = 0;
    for (int i = 0; i < c->count; ++i) {
        const int m = c->m[i], n = c->n[i];
        const size_t mn = (size_t)m * n;
        const double *r = rho + off, *uu = u + off, *vv = v + off, *pp = p + off;
        keep.emplace_back(mn);
        std::vector<double> &E = keep.back();
        for (size_t k = 0; k < mn; ++k) {
            double t = 0.5 * uu[k] * uu[k] + pp[k] / (gamma - 1.0) / r[k];
            t += 0.5 * vv[k] * vv[k];
            E[k] = t;
        }
        const BatchGrid &g = c->grids[i];
        const double *src[5] = {r, uu, vv, pp, E.data()};
        for (int f = 0; f < 5; ++f)
            BCHECK(hipMemcpy2DAsync(g.st[0] + (size_t)f * g.plane, (size_t)g.ld * 8, src[f], (size_t)n * 8,
                                    (size_t)n * 8, (size_t)m, hipMemcpyHostToDevice, c->stream));
        // BC -1 ghosts from the initial data (bound_cond_slope_limiter_x.c:37-46)
        keep.emplace_back((size_t)8 * (m + n));
        std::vector<double> &b = keep.back();
        double *xl = b.data(), *xh = xl + (size_t)n * 4, *yl = xh + (size_t)n * 4, *yh = yl + (size_t)m * 4;
        const double *q4[4] = {r, uu, vv, pp};
        for (int ii = 0; ii < n; ++ii)
            for (int q = 0; q < 4; ++q) {
                xl[(size_t)ii * 4 + q] = q4[q][ii];
                xh[(size_t)ii * 4 + q] = q4[q][(size_t)(m - 1) * n + ii];
            }
        for (int j = 0; j < m; ++j)
            for (int q = 0; q < 4; ++q) {
                yl[(size_t)j * 4 + q] = q4[q][(size_t)j * n];
                yh[(size_t)j * 4 + q] = q4[q][(size_t)j * n + n - 1];
            }
        BCHECK(hipMemcpyAsync((void *)g.bnd[0], b.data(), b.size() * 8, hipMemcpyHostToDevice, c->stream));
        BCHECK(launch_cfl(g.st[0], g.st[0] + g.plane, g.st[0] + 2 * g.plane, g.st[0] + 3 * g.plane, m, n, g.ld, gamma,
                          fmin(c->hx[i], c->hy[i]), c->cfl_bits + i, c->stream));
        off += mn;
    }
    std::vector<unsigned long long> bits((size_t)c->count);
    BCHECK(hipMemcpyAsync(bits.data(), c->cfl_bits, bits.size() * 8, hipMemcpyDeviceToHost, c->stream));
    BCHECK(hipStreamSynchronize(c->stream));
    c->h_S_max.resize(c->count);
    c->sched.clear();
    for (int i = 0; i < c->count; ++i) {
        memcpy(&c->h_S_max[i], &bits[i], 8);
        double cfg[HGRP_N_CONF];
        memcpy(cfg, c->cfg, sizeof cfg);
        cfg[10] = c->hx[i];
        cfg[11] = c->hy[i];
        c->sched.push_back(schedule(cfg, c->h_S_max[i]));
        BatchGrid &g = c->grids[i];
        g.tau = c->sched[i].tau;
        g.tau_last = c->sched[i].tau_last;
        g.last_step = c->sched[i].last_step;
        g.halted = 0;
    }
    BCHECK(hipMemcpy(c->grids_d, c->grids.data(), (size_t)c->count * sizeof(BatchGrid), hipMemcpyHostToDevice));
    // grids by number of sweeps, longest first (ties: index order)
    c->order.resize(c->count);
    for (int i = 0; i < c->count; ++i) c->order[i] = i;
    std::stable_sort(c->order.begin(), c->order.end(),
                     [&](int a, int b) { return c->sched[a].sweeps > c->sched[b].sweeps; });
    for (auto &d : c->tiles_d)
        for (auto *&t : d)
            if (t) { (void)hipFree(t); t = nullptr; }
    c->res_k.assign(c->count, 1);
    c->res_status.assign(c->count, 0);
    c->res_cs.assign(c->count, 0);
    c->res_steps.assign(c->count, 0);
    c->res_time.assign(c->count, 0.0);
    c->uploaded = true;
    return 0;
}

namespace {
// per launch: the argument block of sweep g of the schedule
BatchSweep sweep_args(hgrp2db *c, int g, int slot, int prev_slot, int tiles_dir, int tiles_v)
{
    BatchSweep b = {};
    const int s = g / 3, w = g % 3;
    b.grids = c->grids_d;
    b.tiles = c->tiles_d[tiles_dir][tiles_v];
    b.dir = w == 1 ? 1 : 0;
    b.in = g & 1;
    const int before = b.dir == 0 ? 2 * s + (w > 0 ? 1 : 0) : s;    // earlier sweeps in this direction
    b.sin = before & 1;
    b.limiter3 = before > 0 ? 1 : 0;
    b.step = s;
    b.slot = slot;
    b.prev_slot = prev_slot;
    b.bc = (int)c->cfg[b.dir == 0 ? 17 : 18];
    b.conv = c->newton_rule == HGRP_NEWTON_CONVERGE;
    b.eps = c->cfg[4];
    b.alpha = c->cfg[41];
    b.tol = c->newton_tol > 0.0 ? c->newton_tol : b.eps;
    b.mlp = c->use_mlp ? c->mlp : nullptr;
    b.gas = make_gas(c->cfg[6]);
    return b;
}

int newton_mode_of(const hgrp2db *c)
{
    return c->use_mlp ? NM_LEARNED : c->newton_rule == HGRP_NEWTON_CONVERGE ? NM_CONV : NM_REF;
}

// grids of `order` that have sweep g
int active_prefix(const hgrp2db *c, int g)
{
    int lo = 0, hi = c->count;      // order is sorted by sweeps, descending
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (c->sched[c->order[mid]].sweeps > g) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int enqueue(hgrp2db *c, int g, int slot, int prev_slot)
{
    const int dir = g % 3 == 1 ? 1 : 0, v = dir == 0 ? c->xv : c->yv;
    int rc = build_tiles(c, dir, v);
    if (rc) return rc;
    const int r = active_prefix(c, g);
    const BatchSweep b = sweep_args(c, g, slot, prev_slot, dir, v);
    BCHECK(launch_batch_sweep(b, newton_mode_of(c), v, c->prefix[dir][v][r], c->stream));
    return 0;
}
}  // namespace

extern "C" int hgrp2db_run(hgrp2db *c, double *ms, long long *cell_updates)
{
    if (!c || !c->uploaded) return ARG - 1;
    int G = 0;
    for (const Sched &S : c->sched) G = std::max(G, S.sweeps);
    std::vector<char> ended((size_t)c->count, 0);
    for (int i = 0; i < c->count; ++i) {
        const Sched &S = c->sched[i];
        c->res_k[i] = S.k;
        c->res_status[i] = S.status;
        c->res_cs[i] = S.sweeps & 1;
        c->res_steps[i] = (int)S.t.size() - 1;
        c->res_time[i] = S.time_c;
    }
    BCHECK(hipMemsetAsync(c->status, 0, (size_t)c->count * RING * sizeof(SweepStatus), c->stream));
    std::vector<SweepStatus> st((size_t)c->count * RING);
    bool changed = false;
    // reads the rings of the last `used` sweeps (from g0); settles every grid
    // with a failure the way the one-grid loop does (hgrp2d_solve's settle)
    auto check = [&](int g0, int used) -> int {
        BCHECK(hipMemcpyAsync(st.data(), c->status, st.size() * sizeof(SweepStatus), hipMemcpyDeviceToHost,
                              c->stream));
        BCHECK(hipStreamSynchronize(c->stream));
        for (int i = 0; i < c->count; ++i) {
            if (ended[i]) continue;
            for (int r = 0; r < used; ++r) {
                const SweepStatus &S = st[(size_t)i * RING + r];
                if (!S.fail_key && !(S.flags & 1u)) continue;
                const int g = g0 + r, s = g / 3, w = g % 3;
                const Sched &P = c->sched[i];
                const int code = S.fail_key ? (int)((S.fail_key - 1ull) & 3ull) : 0;
                c->res_k[i] = w == 2 ? s + 2 : s + 1;
                const double tau = s == P.last_step ? P.tau_last : P.tau;
                double time_c = w == 2 ? P.t[s + 1] : P.t[s];
                if (code == 1) {
                    printf("<0.0 or non-finite reconstruction in the %c sweep of step %d (grid %d)\n", w == 1 ? 'y' : 'x',
                           c->res_k[i], i);
                    c->res_cs[i] = g & 1;
                    c->res_status[i] = 1;
                } else {
                    if (code == 2)
                        printf("GRP star/derivative error in the %c sweep of step %d (grid %d)\n", w == 1 ? 'y' : 'x',
                               c->res_k[i], i);
                    if (S.flags & 1u)
                        printf("<0.0 error in the update of the %c sweep of step %d (grid %d)\n", w == 1 ? 'y' : 'x',
                               c->res_k[i], i);
                    c->res_cs[i] = (g + 1) & 1;
                    if (w == 1) time_c += tau;
                    c->res_status[i] = 2;
                }
                c->res_time[i] = time_c;
                c->res_steps[i] = s + (w >= 1 ? 1 : 0);
                ended[i] = 1;
                c->grids[i].halted = 1;
                changed = true;
                break;
            }
        }
        if (changed) {
            BCHECK(hipMemcpyAsync(c->grids_d, c->grids.data(), (size_t)c->count * sizeof(BatchGrid),
                                  hipMemcpyHostToDevice, c->stream));
            changed = false;
        }
        BCHECK(hipMemsetAsync(c->status, 0, (size_t)c->count * RING * sizeof(SweepStatus), c->stream));
        return 0;
    };
    BCHECK(hipEventRecord(c->ev[0], c->stream));
    int g0 = 0;
    for (int g = 0; g < G; ++g) {
        const int slot = g - g0;
        const int rc = enqueue(c, g, slot, slot > 0 ? slot - 1 : -1);
        if (rc) return rc;
        if (slot + 1 == RING || g + 1 == G) {
            const int r = check(g0, slot + 1);
            if (r) return r;
            g0 = g + 1;
        }
    }
    BCHECK(hipEventRecord(c->ev[1], c->stream));
    BCHECK(hipEventSynchronize(c->ev[1]));
    float t = 0.f;
    BCHECK(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    long long cu = 0;
    for (int i = 0; i < c->count; ++i) cu += (long long)c->m[i] * c->n[i] * c->res_steps[i];
    for (int i = 0; i < c->count; ++i) {
        if (c->grids[i].halted) {
            c->grids[i].halted = 0;
        }
    }
    BCHECK(hipMemcpy(c->grids_d, c->grids.data(), (size_t)c->count * sizeof(BatchGrid), hipMemcpyHostToDevice));
    if (ms) *ms = t;
    if (cell_updates) *cell_updates = cu;
    c->uploaded = false;             // the state is the final one: upload again for a new run
    return 0;
}

extern "C" int hgrp2db_results(hgrp2db *c, int *k, int *status, double *time_c, int *steps)
{
    if (!c || c->res_k.size() != (size_t)c->count) return ARG - 1;     // nothing uploaded yet
    for (int i = 0; i < c->count; ++i) {
        if (k) k[i] = c->res_k[i];
        if (status) status[i] = c->res_status[i];
        if (time_c) time_c[i] = c->res_time[i];
        if (steps) steps[i] = c->res_steps[i];
    }
    return 0;
}

extern "C" int hgrp2db_download(hgrp2db *c, double *rho, double *u, double *v, double *p, double *e)
{
    if (!c || c->res_cs.size() != (size_t)c->count) return ARG - 1;    // nothing uploaded yet
    double *dst[5] = {rho, u, v, p, e};
    size_t off = 0;
    for (int i = 0; i < c->count; ++i) {
        const BatchGrid &g = c->grids[i];
        const double *src = g.st[c->res_cs[i]];
        for (int f = 0; f < 5; ++f) {
            if (!dst[f]) continue;
            BCHECK(hipMemcpy2DAsync(dst[f] + off, (size_t)g.n * 8, src + (size_t)f * g.plane, (size_t)g.ld * 8,
                                    (size_t)g.n * 8, (size_t)g.m, hipMemcpyDeviceToHost, c->stream));
        }
        off += (size_t)g.m * g.n;
    }
    BCHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hgrp2db_tune(hgrp2db *c, int steps, int *xv_out, int *yv_out, double *ms_out)
{
    if (!c || !c->uploaded || steps < 1) return ARG - 1;
    int G = 0;
    for (const Sched &S : c->sched) G = std::max(G, S.sweeps);
    const int n_sw = std::min(3 * steps, G);
    if (n_sw < 2) {
        if (xv_out) *xv_out = c->xv;
        if (yv_out) *yv_out = c->yv;
        return 0;
    }
    double *save = nullptr;
    BCHECK(hipMalloc((void **)&save, c->block_doubles * sizeof(double)));
    const int xv0 = c->xv, yv0 = c->yv;
    int rc = 0;
    auto fail = [&](int r) { c->xv = xv0; c->yv = yv0; rc = r; };
    if (hipMemcpyAsync(save, c->block, c->block_doubles * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        fail(-2);
    std::vector<hipEvent_t> ev((size_t)2 * n_sw);
    for (auto &e : ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) fail(-2);
    // device time of the sweeps of one direction over the first n_sw sweeps
    auto trial = [&](int dir, double *t) -> int {
        if (hipMemcpyAsync(c->block, save, c->block_doubles * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return -2;
        for (int g = 0; g < n_sw; ++g) {
            if (hipEventRecord(ev[(size_t)2 * g], c->stream) != hipSuccess) return -2;
            const int r = enqueue(c, g, g % RING, -1);
            if (r) return r;
            if (hipEventRecord(ev[(size_t)2 * g + 1], c->stream) != hipSuccess) return -2;
        }
        if (hipStreamSynchronize(c->stream) != hipSuccess) return -2;
        double sum = 0.0;
        int cnt = 0;
        for (int g = 0; g < n_sw; ++g) {
            if ((g % 3 == 1 ? 1 : 0) != dir) continue;
            float x = 0.f;
            if (hipEventElapsedTime(&x, ev[(size_t)2 * g], ev[(size_t)2 * g + 1]) != hipSuccess) return -2;
            sum += x;
            ++cnt;
        }
        *t = cnt ? sum / cnt : 0.0;
        return 0;
    };
    double best_ms[2] = {0.0, 0.0};
    for (int dir = 0; dir < 2 && !rc; ++dir) {
        const int nv = dir == 0 ? NX_VAR : NY_VAR;
        int pick = dir == 0 ? c->xv : c->yv;
        double best = INFINITY;
        for (int v = 0; v < nv && !rc; ++v) {
            if (dir == 0) c->xv = v; else c->yv = v;
            double t1 = 0.0, t2 = 0.0;
            int r = trial(dir, &t1);
            if (!r) r = trial(dir, &t2);
            if (r) { fail(r); break; }
            const double t = std::min(t1, t2);
            if (t < best) { best = t; pick = v; }
        }
        if (rc) break;
        if (dir == 0) c->xv = pick; else c->yv = pick;
        best_ms[dir] = best;
    }
    // the state the run starts from, and clean status rings
    if (hipMemcpyAsync(c->block, save, c->block_doubles * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(c->status, 0, (size_t)c->count * RING * sizeof(SweepStatus), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        if (!rc) fail(-2);
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(save);
    if (rc) return rc;
    if (xv_out) *xv_out = c->xv;
    if (yv_out) *yv_out = c->yv;
    if (ms_out) { ms_out[0] = best_ms[0]; ms_out[1] = best_ms[1]; }
    return 0;
}
