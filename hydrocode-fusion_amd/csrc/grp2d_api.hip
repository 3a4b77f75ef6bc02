// grp2d_api.hip -- host side of the 2D split GRP solver: device buffers, the
// reference's time loop (schedule, frozen tau, plot frames, stop/abort rules)
// and the batched interface entry points.  C ABI in include/hydrocode_grp.h.
//
// The time loop restates GRP_solver_2D_split_EUL_source
// (src/finite_volume/grp_solver_2D_split_EUL_source.c:147-341), or
// GRP_solver_2D_EUL_source (src/finite_volume/grp_solver_2D_EUL_source.c:147-300)
// when config[33] is 0, on the host side of a GPU pipeline: sweeps (steps) are
// enqueued without waiting; each records its failures in a status slot and one
// whose predecessor failed does nothing, so the host can check every few dozen
// launches and still end exactly where the reference ends.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <vector>

#include "../../include/hydrocode_grp.h"
#include "grp_device.hpp"
#include "sweep2d.hpp"
#include "unsplit2d.hpp"

namespace hgrp {
hipError_t launch_sweep_x(const SweepArgs &a, int variant, hipStream_t st);
hipError_t launch_sweep_y(const SweepArgs &a, int variant, hipStream_t st);
const char *sweep_kernel_name(int dir, int variant, int nm);
int sweep_variants(int dir);
hipError_t launch_sweep_x_tf(const SweepArgs &a, hipStream_t st);
const char *sweep_tf_kernel_name(int dir, int nm);
hipError_t launch_sweep_y_tf(const SweepArgs &a, hipStream_t st);
hipError_t launch_cfl_tf(const double *rho, const double *u, const double *v, const double *p, const double *fa,
                         int m, int n, int ld, const TwoGas &two, double hmin, unsigned long long *out, hipStream_t st);
hipError_t launch_energy_tf(const double *rho, const double *u, const double *v, const double *p, const double *fa,
                            double *e, int m, int n, int ld, const TwoGas &two, hipStream_t st);
hipError_t launch_cfl(const double *rho, const double *u, const double *v, const double *p, int m, int n,
                      int ld, double gamma, double hmin, unsigned long long *out, hipStream_t st);
hipError_t launch_unsplit_step(const UnsplitArgs &a, hipStream_t st);
const char *unsplit_kernel_name(int m, int n, int nm, int g2d, int form);
int unsplit_max_rows();
}

using namespace hgrp;

#define HCHECK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "hydrogrp: %s failed: %s (%s:%d)\n", #x,  \
                    hipGetErrorString(e_), __FILE__, __LINE__);       \
            return -(int)e_ - 1;                                      \
        }                                                             \
    } while (0)

static constexpr int STATUS_SLOTS = 96;    // sweeps between host checks

// Where the reference's time loop stands (hgrp2d_solve's locals), for
// checkpoints taken in a plot-frame callback and for resuming from them.
struct LoopPos {
    int valid, k, DS, nt, nt_plot;
    double time_c, tau, half_tau, half_nu, mu;
};
static constexpr int ARG_ERROR = -1000;

struct hgrp2d {
    int m, n, ld;
    double cfg[HGRP_N_CONF];
    hipStream_t stream;
    double *block;              // one device allocation for everything below
    double *st[2][5];           // state ping-pong: rho u v p e
    double *sx[2][4], *sy[2][4];
    double *bnd[4];             // BC -1 ghosts from frame 0: x_lo, x_hi (n*4), y_lo, y_hi (m*4)
    SweepStatus *status;        // STATUS_SLOTS slots
    unsigned long long *cfl_bits;
    int cs, cx, cy;             // current buffer of state / x-slopes / y-slopes
    int found_x, found_y;       // the reference's find_bound_x / find_bound_y
    double h_S_max;
    bool uploaded;
    LoopPos at;                 // inside a frame callback: where the loop goes on after it
    LoopPos resume;             // set by hgrp2d_restore: the next solve starts here
    double newton_tol;          // <= 0: the reference's (eps)
    int newton_rule;            // HGRP_NEWTON_REFERENCE or HGRP_NEWTON_CONVERGE
    int xv, yv;                 // launch variants of the X and Y sweeps (hgrp2d_tune)
    int g2d;                    // the unsplit step's interface solver: 0 Q1D, 1 G2D
    int uform;                  // the unsplit step's launch form: 0 by grid size, 1 tile, 2 strip
    MlpPacked *mlp;             // device copy of the learned-guess weights
    bool use_mlp;
    bool two_fluid;             // volume fraction of fluid a carried (hgrp2d_set_fluids)
    double *tf_block;           // fa[2], a-slopes in x [2] and in y [2]
    double *fa[2], *sax[2], *say[2];
    TwoGas two;
    hipEvent_t ev[2];
};

static size_t pitch_of(int n) { return (size_t)((n + 31) / 32) * 32; }

extern "C" int hgrp2d_create(hgrp2d **out, int m, int n, const double *config)
{
    if (!out || m < 1 || n < 1 || !config) return ARG_ERROR - 1;
    *out = nullptr;
    hgrp2d *c = (hgrp2d *)calloc(1, sizeof(hgrp2d));
    if (!c) return ARG_ERROR - 5;
    c->m = m; c->n = n; c->ld = (int)pitch_of(n);
    memcpy(c->cfg, config, sizeof c->cfg);
    const size_t cells = (size_t)m * c->ld;
    const size_t nb = (size_t)(2 * 5 + 2 * 4 + 2 * 4) * cells + 2 * (size_t)n * 4 + 2 * (size_t)m * 4;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&c->block, nb * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&c->status, STATUS_SLOTS * sizeof(SweepStatus) + 64);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[1]);
    if (e != hipSuccess) {
        fprintf(stderr, "hydrogrp: device allocation of %.1f MiB failed: %s\n",
                nb * 8.0 / 1048576.0, hipGetErrorString(e));
        if (c->block) (void)hipFree(c->block);
        if (c->status) (void)hipFree(c->status);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        free(c);
        return -(int)e - 1;
    }
    c->cfl_bits = (unsigned long long *)((char *)c->status + STATUS_SLOTS * sizeof(SweepStatus));
    double *q = c->block;
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f) { c->st[b][f] = q; q += cells; }
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 4; ++f) { c->sx[b][f] = q; q += cells; }
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 4; ++f) { c->sy[b][f] = q; q += cells; }
    c->bnd[0] = q; q += (size_t)n * 4;
    c->bnd[1] = q; q += (size_t)n * 4;
    c->bnd[2] = q; q += (size_t)m * 4;
    c->bnd[3] = q;
    // slopes start at zero (only read after the first sweep of a direction has written them)
    if (hipMemsetAsync(c->block, 0, nb * sizeof(double), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        hgrp2d_destroy(c);
        return -1;
    }
    *out = c;
    return 0;
}

extern "C" int hgrp2d_destroy(hgrp2d *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->block) (void)hipFree(c->block);
    if (c->status) (void)hipFree(c->status);
    if (c->mlp) (void)hipFree(c->mlp);
    if (c->tf_block) (void)hipFree(c->tf_block);
    if (c->ev[0]) (void)hipEventDestroy(c->ev[0]);
    if (c->ev[1]) (void)hipEventDestroy(c->ev[1]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c);
    return 0;
}

static int copy_rows_in(hgrp2d *c, double *dst, const double *src)
{
    HCHECK(hipMemcpy2DAsync(dst, (size_t)c->ld * 8, src, (size_t)c->n * 8, (size_t)c->n * 8, (size_t)c->m,
                            hipMemcpyHostToDevice, c->stream));
    return 0;
}

extern "C" int hgrp2d_upload(hgrp2d *c, const double *rho, const double *u, const double *v,
                             const double *p, const double *e)
{
    if (!c || !rho || !u || !v || !p) return ARG_ERROR - 1;
    c->resume.valid = 0;                  // fresh initial data: the loop starts at t = 0
    c->two_fluid = false;                 // a new state needs its volume fraction again
    const int m = c->m, n = c->n;
    const double gamma = c->cfg[6];
    const size_t mn = (size_t)m * n;
    std::vector<double> E;
    if (!e) {
        E.resize(mn);
        for (size_t k = 0; k < mn; ++k) {
            double t = 0.5 * u[k] * u[k] + p[k] / (gamma - 1.0) / rho[k];
            t += 0.5 * v[k] * v[k];
            E[k] = t;
        }
        e = E.data();
    }
    const double *src[5] = {rho, u, v, p, e};
    for (int f = 0; f < 5; ++f) {
        int rc = copy_rows_in(c, c->st[0][f], src[f]);
        if (rc) return rc;
    }
    // BC -1 ghosts come from the initial data (bound_cond_slope_limiter_x.c:37-46)
    std::vector<double> b((size_t)2 * n * 4 + (size_t)2 * m * 4);
    double *xl = b.data(), *xh = xl + (size_t)n * 4, *yl = xh + (size_t)n * 4, *yh = yl + (size_t)m * 4;
    const double *q4[4] = {rho, u, v, p};
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < 4; ++q) {
            xl[(size_t)i * 4 + q] = q4[q][i];
            xh[(size_t)i * 4 + q] = q4[q][(size_t)(m - 1) * n + i];
        }
    for (int j = 0; j < m; ++j)
        for (int q = 0; q < 4; ++q) {
            yl[(size_t)j * 4 + q] = q4[q][(size_t)j * n];
            yh[(size_t)j * 4 + q] = q4[q][(size_t)j * n + n - 1];
        }
    HCHECK(hipMemcpyAsync(c->bnd[0], b.data(), b.size() * 8, hipMemcpyHostToDevice, c->stream));
    // frozen CFL reduction over the initial data (grp_solver_2D_split_EUL_source.c:183-191)
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    HCHECK(hipMemcpyAsync(c->cfl_bits, &inf_bits, 8, hipMemcpyHostToDevice, c->stream));
    HCHECK(launch_cfl(c->st[0][0], c->st[0][1], c->st[0][2], c->st[0][3], m, n, c->ld, gamma,
                      fmin(c->cfg[10], c->cfg[11]), c->cfl_bits, c->stream));
    unsigned long long bits = 0;
    HCHECK(hipMemcpyAsync(&bits, c->cfl_bits, 8, hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    memcpy(&c->h_S_max, &bits, 8);
    c->cs = c->cx = c->cy = 0;
    c->found_x = c->found_y = 0;
    c->uploaded = true;
    return 0;
}

extern "C" int hgrp2d_download(hgrp2d *c, double *rho, double *u, double *v, double *p, double *e)
{
    if (!c) return ARG_ERROR - 1;
    double *dst[5] = {rho, u, v, p, e};
    for (int f = 0; f < 5; ++f) {
        if (!dst[f]) continue;
        HCHECK(hipMemcpy2DAsync(dst[f], (size_t)c->n * 8, c->st[c->cs][f], (size_t)c->ld * 8, (size_t)c->n * 8,
                                (size_t)c->m, hipMemcpyDeviceToHost, c->stream));
    }
    HCHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hgrp2d_h_over_smax(hgrp2d *c, double *h)
{
    if (!c || !h || !c->uploaded) return ARG_ERROR - 1;
    *h = c->h_S_max;
    return 0;
}

extern "C" const char *hgrp2d_sweep_kernel(hgrp2d *c, int dir)
{
    if (!c || dir < 0 || dir > 2) return "";
    const int nm = c->use_mlp ? NM_LEARNED : c->newton_rule == HGRP_NEWTON_CONVERGE ? NM_CONV : NM_REF;
    if (dir == 2) return unsplit_kernel_name(c->m, c->n, nm, c->g2d, c->uform);
    if (c->two_fluid) return sweep_tf_kernel_name(dir, nm);
    return sweep_kernel_name(dir, dir == 0 ? c->xv : c->yv, nm);
}

// h/S_max of the current state: the reference's CFL loop reads CV[0], which is
// the live state until the first plot frame has been copied off
// (grp_solver_2D_split_EUL_source.c:183-191 with nt == 0)
static int refresh_cfl(hgrp2d *c)
{
    const double hmin = fmin(c->cfg[10], c->cfg[11]);
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    double *const *st = c->st[c->cs];
    HCHECK(hipMemcpyAsync(c->cfl_bits, &inf_bits, 8, hipMemcpyHostToDevice, c->stream));
    if (c->two_fluid)
        HCHECK(launch_cfl_tf(st[0], st[1], st[2], st[3], c->fa[c->cs], c->m, c->n, c->ld, c->two, hmin, c->cfl_bits,
                             c->stream));
    else
        HCHECK(launch_cfl(st[0], st[1], st[2], st[3], c->m, c->n, c->ld, c->cfg[6], hmin, c->cfl_bits, c->stream));
    unsigned long long bits = 0;
    HCHECK(hipMemcpyAsync(&bits, c->cfl_bits, 8, hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    memcpy(&c->h_S_max, &bits, 8);
    return 0;
}

extern "C" int hgrp2d_set_fluids(hgrp2d *c, double gamma_b, const double *frac_a)
{
    if (!c || !c->uploaded || !frac_a || !(gamma_b > 1.0)) return ARG_ERROR - 1;
    for (int d = 17; d <= 18; ++d) {
        const int b = (int)c->cfg[d];
        if (b != -2 && b != -4 && b != -7 && b != -24) {
            printf("The two-fluid sweeps take boundary conditions -2, -4, -7 and -24 only (config[%d] = %d).\n", d, b);
            return ARG_ERROR - 4;
        }
    }
    const size_t cells = (size_t)c->m * c->ld;
    if (!c->tf_block) {
        HCHECK(hipMalloc((void **)&c->tf_block, 6 * cells * sizeof(double)));
        for (int b = 0; b < 2; ++b) {
            c->fa[b] = c->tf_block + b * cells;
            c->sax[b] = c->tf_block + (2 + b) * cells;
            c->say[b] = c->tf_block + (4 + b) * cells;
        }
    }
    HCHECK(hipMemsetAsync(c->tf_block, 0, 6 * cells * sizeof(double), c->stream));
    HCHECK(hipMemcpy2DAsync(c->fa[c->cs], (size_t)c->ld * 8, frac_a, (size_t)c->n * 8, (size_t)c->n * 8,
                            (size_t)c->m, hipMemcpyHostToDevice, c->stream));
    c->two.a = make_gas(c->cfg[6]);
    c->two.b = make_gas(gamma_b);
    HCHECK(launch_energy_tf(c->st[c->cs][0], c->st[c->cs][1], c->st[c->cs][2], c->st[c->cs][3], c->fa[c->cs],
                            c->st[c->cs][4], c->m, c->n, c->ld, c->two, c->stream));
    // the frozen CFL bound with each cell's own gamma
    const double hmin = fmin(c->cfg[10], c->cfg[11]);
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    HCHECK(hipMemcpyAsync(c->cfl_bits, &inf_bits, 8, hipMemcpyHostToDevice, c->stream));
    HCHECK(launch_cfl_tf(c->st[c->cs][0], c->st[c->cs][1], c->st[c->cs][2], c->st[c->cs][3], c->fa[c->cs], c->m, c->n,
                         c->ld, c->two, hmin, c->cfl_bits, c->stream));
    unsigned long long bits = 0;
    HCHECK(hipMemcpyAsync(&bits, c->cfl_bits, 8, hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    memcpy(&c->h_S_max, &bits, 8);
    c->two_fluid = true;
    return 0;
}

extern "C" int hgrp2d_download_fraction(hgrp2d *c, double *frac_a)
{
    if (!c || !frac_a || !c->two_fluid) return ARG_ERROR - 1;
    HCHECK(hipMemcpy2DAsync(frac_a, (size_t)c->n * 8, c->fa[c->cs], (size_t)c->ld * 8, (size_t)c->n * 8, (size_t)c->m,
                            hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hgrp2d_set_newton(hgrp2d *c, double tol, int rule, const float *params)
{
    if (!c || !(tol == tol) || (rule != HGRP_NEWTON_REFERENCE && rule != HGRP_NEWTON_CONVERGE)) return ARG_ERROR - 1;
    // the learned guess runs under the converge rule: under the reference's a
    // solve can stop on a different iterate than from the reference's guess
    if (params && rule != HGRP_NEWTON_CONVERGE) return ARG_ERROR - 2;
    HCHECK(hipStreamSynchronize(c->stream));
    c->newton_tol = tol;
    c->newton_rule = rule;
    c->use_mlp = false;
    if (params) {
        MlpPacked h;
        mlp_pack_host(params, &h);
        if (!c->mlp) HCHECK(hipMalloc((void **)&c->mlp, sizeof(MlpPacked)));
        HCHECK(hipMemcpy(c->mlp, &h, sizeof h, hipMemcpyHostToDevice));
        c->use_mlp = true;
    }
    return 0;
}

extern "C" int hgrp2d_set_interface_solver(hgrp2d *c, int solver)
{
    if (!c || (solver != HGRP_SOLVER_Q1D && solver != HGRP_SOLVER_G2D)) return ARG_ERROR - 1;
    // the split sweeps have no tangential derivatives, where G2D is Q1D but
    // for its distance (which includes v) and the rounding of a few products;
    // they keep the reference's solver
    if (solver == HGRP_SOLVER_G2D && c->cfg[33] != 0.0) {
        fprintf(stderr, "hydrogrp: the G2D interface solver runs in the unsplit scheme (config[33] = 0) only\n");
        return ARG_ERROR - 2;
    }
    HCHECK(hipStreamSynchronize(c->stream));
    c->g2d = solver == HGRP_SOLVER_G2D;
    return 0;
}

// One enqueued sweep and the host loop state it belongs to.
struct SweepRec {
    int dir;              // 0 x, 1 y
    int k, DS;            // loop counter and pass kind when it was enqueued
    double time_c;        // time_c when it was enqueued
    double tau;           // the step's tau
    int in_s, in_d;       // state buffer and slope buffer read
};

static void fill_args(hgrp2d *c, SweepArgs &a, int dir, double tau_flux, double nu, int slot, int prev_slot)
{
    const int in = c->cs, out = 1 - c->cs;
    const int sin = dir == 0 ? c->cx : c->cy;
    // the planes of each group are consecutive in c->block (hgrp2d_create)
    a.st = c->st[in][0];
    a.st_o = c->st[out][0];
    a.s = dir == 0 ? c->sx[sin][0] : c->sy[sin][0];
    a.s_o = dir == 0 ? c->sx[1 - sin][0] : c->sy[1 - sin][0];
    a.plane = (size_t)c->m * c->ld;
    a.bnd_lo = c->bnd[dir == 0 ? 0 : 2];
    a.bnd_hi = c->bnd[dir == 0 ? 1 : 3];
    a.m = c->m; a.n = c->n; a.ld = c->ld;
    a.bc = (int)c->cfg[dir == 0 ? 17 : 18];
    a.limiter3 = dir == 0 ? c->found_x : c->found_y;
    a.h = c->cfg[dir == 0 ? 10 : 11];
    a.hh = 0.5 * a.h;
    a.div_h = recip_div(a.h);
    a.tau = tau_flux;
    a.nu = nu;
    a.gas = make_gas(c->cfg[6]);
    a.eps = c->cfg[4];
    a.alpha = c->cfg[41];
    a.tol = c->newton_tol > 0.0 ? c->newton_tol : a.eps;
    a.conv = c->newton_rule == HGRP_NEWTON_CONVERGE;
    if (c->two_fluid) {
        a.fa = c->fa[in]; a.fa_o = c->fa[out];
        a.sa = dir == 0 ? c->sax[sin] : c->say[sin];
        a.sa_o = dir == 0 ? c->sax[1 - sin] : c->say[1 - sin];
        a.two = c->two;
    }
    a.mlp = c->use_mlp ? c->mlp : nullptr;
    a.status = c->status + slot;
    a.prev = prev_slot >= 0 ? c->status + prev_slot : nullptr;
}

static int enqueue_sweep(hgrp2d *c, int dir, double tau_flux, double nu, int slot, int prev_slot)
{
    SweepArgs a;
    fill_args(c, a, dir, tau_flux, nu, slot, prev_slot);
    if (c->two_fluid)
        HCHECK(dir == 0 ? launch_sweep_x_tf(a, c->stream) : launch_sweep_y_tf(a, c->stream));
    else
        HCHECK(dir == 0 ? launch_sweep_x(a, c->xv, c->stream) : launch_sweep_y(a, c->yv, c->stream));
    c->cs = 1 - c->cs;
    if (dir == 0) { c->cx = 1 - c->cx; c->found_x = 1; }
    else          { c->cy = 1 - c->cy; c->found_y = 1; }
    return 0;
}

static bool valid_bc(int b) { return b == -1 || b == -2 || b == -4 || b == -7 || b == -24; }

extern "C" int hgrp2d_bench(hgrp2d *c, int steps, int sample_every, double *ms_total, double *ms_x, double *ms_y)
{
    if (!c || !c->uploaded || steps < 1) return ARG_ERROR - 1;
    if (!valid_bc((int)c->cfg[17]) || !valid_bc((int)c->cfg[18])) return ARG_ERROR - 4;
    const double tau = c->cfg[7] * c->h_S_max;
    const double half_tau = tau * 0.5, half_nu = half_tau / c->cfg[10], mu = tau / c->cfg[11];
    HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));
    // per-sweep events for the average launch duration of each sweep kind, on
    // every EVERY-th step only (a uniform sample of the run); none with
    // sample_every <= 0, so that the timed stream carries the sweeps alone
    const int every = sample_every > 0 ? sample_every : steps + 1;
    const int ns = sample_every > 0 ? (steps + every - 1) / every : 0;
    std::vector<hipEvent_t> evs((size_t)6 * ns);
    for (auto &e : evs) HCHECK(hipEventCreate(&e));
    HCHECK(hipEventRecord(c->ev[0], c->stream));
    int slot = 0;
    for (int s = 0; s < steps; ++s) {
        const int dirs[3] = {0, 1, 0};
        const bool sampled = ns > 0 && s % every == 0;
        const size_t e0 = (size_t)6 * (s / every);
        for (int w = 0; w < 3; ++w) {
            const int d = dirs[w];
            if (sampled) HCHECK(hipEventRecord(evs[e0 + 2 * w], c->stream));
            // status slots are not chained here: the benchmark must run every sweep
            int rc = enqueue_sweep(c, d, d == 0 ? half_tau : tau, d == 0 ? half_nu : mu, slot % STATUS_SLOTS, -1);
            if (rc) return rc;
            if (sampled) HCHECK(hipEventRecord(evs[e0 + 2 * w + 1], c->stream));
            ++slot;
        }
    }
    HCHECK(hipEventRecord(c->ev[1], c->stream));
    HCHECK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (ms_total) *ms_total = ms;
    double sx = 0, sy = 0;
    for (int k = 0; k < ns; ++k)
        for (int w = 0; w < 3; ++w) {
            float t = 0.f;
            HCHECK(hipEventElapsedTime(&t, evs[(size_t)6 * k + 2 * w], evs[(size_t)6 * k + 2 * w + 1]));
            if (w == 1) sy += t; else sx += t;
        }
    if (ms_x) *ms_x = ns ? sx / (2.0 * ns) : NAN;
    if (ms_y) *ms_y = ns ? sy / ns : NAN;
    for (auto &e : evs) (void)hipEventDestroy(e);
    return 0;
}

extern "C" int hgrp2d_set_unsplit_form(hgrp2d *c, int form)
{
    if (!c || form < HGRP_UNSPLIT_AUTO || form > HGRP_UNSPLIT_STRIP) return ARG_ERROR - 1;
    c->uform = form;
    return 0;
}

static int tune_unsplit(hgrp2d *c, int steps, int *form_out, double *ms_out);

extern "C" int hgrp2d_set_sweeps(hgrp2d *c, int xv, int yv)
{
    if (!c || xv < 0 || xv >= sweep_variants(0) || yv < 0 || yv >= sweep_variants(1)) return ARG_ERROR - 1;
    c->xv = xv;
    c->yv = yv;
    return 0;
}

// The sweep autotuner.  Times every launch variant of the X sweep, then of the
// Y sweep (with the X sweep's best), over `steps` steps from the context's
// current state, two trials each, and keeps the fastest (per-launch device
// time from HIP events around each launch of the kind being tuned).  The
// state, slopes and limiter switches are copied aside on the device first and
// put back after every trial and at the end, so tuning changes no result.
extern "C" int hgrp2d_tune(hgrp2d *c, int steps, int *xv_out, int *yv_out, double *ms_out)
{
    if (!c || !c->uploaded || steps < 1) return ARG_ERROR - 1;
    if (c->two_fluid) return ARG_ERROR - 4;      // the two-fluid sweeps have one launch form
    if (c->cfg[33] == 0.0) {                     // the unsplit scheme: its two launch forms
        if (yv_out) *yv_out = -1;
        if (ms_out) ms_out[1] = NAN;
        return tune_unsplit(c, steps, xv_out, ms_out);
    }
    if (!valid_bc((int)c->cfg[17]) || !valid_bc((int)c->cfg[18])) return ARG_ERROR - 4;
    const size_t cells = (size_t)c->m * c->ld;
    const size_t nb = (size_t)(2 * 5 + 2 * 4 + 2 * 4) * cells;     // state and slope planes
    double *save = nullptr;
    HCHECK(hipMalloc((void **)&save, nb * sizeof(double)));
    if (hipMemcpyAsync(save, c->block, nb * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        (void)hipFree(save);
        return ARG_ERROR - 5;
    }
    const int cs = c->cs, cx = c->cx, cy = c->cy, fx = c->found_x, fy = c->found_y;
    const int xv0 = c->xv, yv0 = c->yv;
    int rc = 0;
    auto put_back = [&]() -> int {
        HCHECK(hipMemcpyAsync(c->block, save, nb * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        c->cs = cs; c->cx = cx; c->cy = cy; c->found_x = fx; c->found_y = fy;
        return 0;
    };
    const double tau = c->cfg[7] * c->h_S_max;
    const double half_tau = tau * 0.5, half_nu = half_tau / c->cfg[10], mu = tau / c->cfg[11];
    std::vector<hipEvent_t> evs((size_t)4 * steps);
    for (auto &e : evs)
        if (hipEventCreate(&e) != hipSuccess) { (void)hipFree(save); return ARG_ERROR - 5; }
    // average device time of one launch of the kind `dir` over `steps` steps
    auto trial = [&](int dir, double *ms) -> int {
        if (int r = put_back()) return r;
        HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));
        int slot = 0, ne = 0;
        for (int s = 0; s < steps; ++s) {
            const int dirs[3] = {0, 1, 0};
            for (int w = 0; w < 3; ++w) {
                const int d = dirs[w];
                if (d == dir) HCHECK(hipEventRecord(evs[(size_t)ne++], c->stream));
                if (int r = enqueue_sweep(c, d, d == 0 ? half_tau : tau, d == 0 ? half_nu : mu, slot++ % STATUS_SLOTS, -1))
                    return r;
                if (d == dir) HCHECK(hipEventRecord(evs[(size_t)ne++], c->stream));
            }
        }
        HCHECK(hipEventSynchronize(evs[(size_t)ne - 1]));
        double sum = 0.0;
        for (int k = 0; k < ne; k += 2) {
            float t = 0.f;
            HCHECK(hipEventElapsedTime(&t, evs[(size_t)k], evs[(size_t)k + 1]));
            sum += t;
        }
        *ms = sum / (ne / 2);
        return 0;
    };
    double best[2] = {0.0, 0.0};
    for (int dir = 0; dir < 2 && !rc; ++dir) {
        int pick = 0;
        for (int v = 0; v < sweep_variants(dir) && !rc; ++v) {
            if (dir == 0) c->xv = v; else c->yv = v;
            double t = 1e300;
            for (int k = 0; k < 2 && !rc; ++k) {
                double ms = 0.0;
                rc = trial(dir, &ms);
                if (!rc && ms < t) t = ms;
            }
            if (!rc && (v == 0 || t < best[dir])) { best[dir] = t; pick = v; }
        }
        if (dir == 0) c->xv = pick; else c->yv = pick;
    }
    if (rc) { c->xv = xv0; c->yv = yv0; }
    const int rb = put_back();
    if (!rc) rc = rb;
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = -(int)e - 1;
    }
    for (auto &e : evs) (void)hipEventDestroy(e);
    (void)hipFree(save);
    if (rc) return rc;
    if (xv_out) *xv_out = c->xv;
    if (yv_out) *yv_out = c->yv;
    if (ms_out) { ms_out[0] = best[0]; ms_out[1] = best[1]; }
    return 0;
}

// ---- checkpoints --------------------------------------------------------------
// Layout: CkptHead, then state rho u v p E, x-slopes, y-slopes (each m*n
// doubles, j outer), then the BC -1 ghost tables (2n*4 + 2m*4 doubles).
struct CkptHead {
    char magic[8];
    int version, m, n, found_x, found_y, pad;
    LoopPos pos;
    double h_S_max;
    double cfg[HGRP_N_CONF];
};
static const char CKPT_MAGIC[8] = {'H', 'G', 'R', 'P', 'C', 'K', '2', 'D'};

static size_t ckpt_bytes(const hgrp2d *c)
{
    const size_t cells = (size_t)c->m * c->n;
    return sizeof(CkptHead) + (13 * cells + 8 * ((size_t)c->m + c->n)) * sizeof(double);
}

extern "C" int hgrp2d_checkpoint_size(hgrp2d *c, size_t *bytes)
{
    if (!c || !bytes) return ARG_ERROR - 1;
    *bytes = ckpt_bytes(c);
    return 0;
}

extern "C" int hgrp2d_checkpoint(hgrp2d *c, void *buf, size_t bytes)
{
    if (!c || !buf || !c->uploaded || bytes < ckpt_bytes(c) || c->two_fluid) return ARG_ERROR - 1;
    CkptHead *h = (CkptHead *)buf;
    memset(h, 0, sizeof *h);
    memcpy(h->magic, CKPT_MAGIC, 8);
    h->version = 1; h->m = c->m; h->n = c->n; h->found_x = c->found_x; h->found_y = c->found_y;
    h->pos = c->at;                       // valid only inside a frame callback
    h->h_S_max = c->h_S_max;
    memcpy(h->cfg, c->cfg, sizeof h->cfg);
    double *d = (double *)(h + 1);
    const size_t cells = (size_t)c->m * c->n, row = (size_t)c->n * 8, pitch = (size_t)c->ld * 8;
    HCHECK(hipStreamSynchronize(c->stream));
    double *planes[13];
    for (int q = 0; q < 5; ++q) planes[q] = c->st[c->cs][q];
    for (int q = 0; q < 4; ++q) { planes[5 + q] = c->sx[c->cx][q]; planes[9 + q] = c->sy[c->cy][q]; }
    for (int q = 0; q < 13; ++q)
        HCHECK(hipMemcpy2DAsync(d + q * cells, row, planes[q], pitch, row, (size_t)c->m, hipMemcpyDeviceToHost,
                                c->stream));
    double *b = d + 13 * cells;
    HCHECK(hipMemcpyAsync(b, c->bnd[0], (size_t)(8 * c->n + 8 * c->m) * 8, hipMemcpyDeviceToHost, c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int hgrp2d_restore(hgrp2d *c, const void *buf, size_t bytes)
{
    if (!c || !buf || bytes < sizeof(CkptHead)) return ARG_ERROR - 1;
    const CkptHead *h = (const CkptHead *)buf;
    if (memcmp(h->magic, CKPT_MAGIC, 8) != 0 || h->version != 1 || h->m != c->m || h->n != c->n ||
        bytes < ckpt_bytes(c))
        return ARG_ERROR - 2;
    const double *d = (const double *)(h + 1);
    const size_t cells = (size_t)c->m * c->n, row = (size_t)c->n * 8, pitch = (size_t)c->ld * 8;
    HCHECK(hipStreamSynchronize(c->stream));
    double *planes[13];
    for (int q = 0; q < 5; ++q) planes[q] = c->st[c->cs][q];
    for (int q = 0; q < 4; ++q) { planes[5 + q] = c->sx[c->cx][q]; planes[9 + q] = c->sy[c->cy][q]; }
    for (int q = 0; q < 13; ++q)
        HCHECK(hipMemcpy2DAsync(planes[q], pitch, d + q * cells, row, row, (size_t)c->m, hipMemcpyHostToDevice,
                                c->stream));
    HCHECK(hipMemcpyAsync(c->bnd[0], d + 13 * cells, (size_t)(8 * c->n + 8 * c->m) * 8, hipMemcpyHostToDevice,
                          c->stream));
    HCHECK(hipStreamSynchronize(c->stream));
    c->found_x = h->found_x;
    c->found_y = h->found_y;
    c->h_S_max = h->h_S_max;              // the frozen CFL bound of the original initial data
    c->resume = h->pos;
    c->uploaded = true;
    return 0;
}

// ---- the reference's time loop -------------------------------------------------
namespace {
struct Loop {
    hgrp2d *c;
    std::vector<SweepRec> recs;      // sweeps since the last check; recs[r] uses status slot r
    int sweeps = 0;

    // reads the status of the pending sweeps; returns the index of the first
    // failed one, -1 if none, or a negative HIP error - 2
    int check(unsigned long long *key, unsigned *flags)
    {
        if (recs.empty()) return -1;
        std::vector<SweepStatus> st(recs.size());
        if (hipMemcpyAsync(st.data(), c->status, st.size() * sizeof(SweepStatus), hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return -2;
        for (size_t r = 0; r < st.size(); ++r)
            if (st[r].fail_key || (st[r].flags & 1u)) {
                *key = st[r].fail_key;
                *flags = st[r].flags;
                return (int)r;
            }
        return -1;
    }
};
}  // namespace

static int solve_unsplit(hgrp2d *c, const double *time_plot, int N_plot, int N_T, hgrp2d_frame_fn fn, void *user,
                         hgrp2d_result *res);

extern "C" int hgrp2d_solve(hgrp2d *c, const double *time_plot, int N_plot, int N_T,
                            hgrp2d_frame_fn fn, void *user, hgrp2d_result *res)
{
    if (!c || !c->uploaded || !time_plot || N_plot < 2 || N_T < 1 || !res) return ARG_ERROR - 1;
    if (c->cfg[33] == 0.0)                 // hydrocode.c:264-267: (_Bool)config[33] picks the split solver
        return solve_unsplit(c, time_plot, N_plot, N_T, fn, user, res);
    const double t_all = c->cfg[1], eps = c->cfg[4], CFL = c->cfg[7];
    const int N = (int)c->cfg[5];
    const double h_x = c->cfg[10], h_y = c->cfg[11];
    double tau = c->cfg[16];
    const int bc_x = (int)c->cfg[17], bc_y = (int)c->cfg[18];
    const auto t_start = std::chrono::steady_clock::now();

    Loop L;
    L.c = c;
    int k = 1, nt = 0, nt_plot = 0, DS = 1, status = 0;
    double time_c = 0.0, half_tau = 0.0, half_nu = 0.0, mu = 0.0;
    bool stop_t = false, skip_frame = false;
    if (c->resume.valid) {                  // continue where a checkpoint was taken
        const LoopPos &r = c->resume;
        k = r.k; DS = r.DS; nt = r.nt; nt_plot = r.nt_plot;
        time_c = r.time_c; tau = r.tau; half_tau = r.half_tau; half_nu = r.half_nu; mu = r.mu;
        skip_frame = true;                  // that iteration's frame was handled before the checkpoint
        c->resume.valid = 0;
        if (nt > N_T - 1) nt = N_T - 1;
    }
    HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));

    // finish at a sweep that failed on the device: restore the loop state it
    // ran in and apply the reference's rule for that failure
    auto settle = [&](int r, unsigned long long key, unsigned flags) {
        const SweepRec &R = L.recs[(size_t)r];
        const int code = key ? (int)((key - 1ull) & 3ull) : 0;
        k = R.k; DS = R.DS; time_c = R.time_c; tau = R.tau;
        L.sweeps += r + 1 - (int)L.recs.size();
        if (code == 1) {                       // ifvar_check failed: goto return_NULL, no update
            printf("<0.0 or non-finite reconstruction in the %c sweep of step %d\n", R.dir ? 'y' : 'x', k);
            c->cs = R.in_s;
            status = 1;
        } else {                               // GRP or positivity failure: update kept, stop
            if (code == 2) printf("GRP star/derivative error in the %c sweep of step %d\n", R.dir ? 'y' : 'x', k);
            if (flags & 1u) printf("<0.0 error in the update of the %c sweep of step %d\n", R.dir ? 'y' : 'x', k);
            c->cs = 1 - R.in_s;
            if (R.dir == 1) time_c += R.tau;
            status = 2;
        }
        (void)flags;
    };
    auto flush = [&]() -> int {               // returns 1 if a failure ended the run
        unsigned long long key = 0;
        unsigned flags = 0;
        const int r = L.check(&key, &flags);
        if (r <= -2) return -2;
        if (r >= 0) { settle(r, key, flags); return 1; }
        L.recs.clear();
        return 0;
    };
    // after a clean check every status slot is still zero, so slots restart at 0
    auto sweep = [&](int dir, double tf, double nu) -> int {
        if ((int)L.recs.size() + 1 > STATUS_SLOTS) {
            const int f = flush();
            if (f) return f;
        }
        const int slot = (int)L.recs.size();
        const int prev = L.recs.empty() ? -1 : slot - 1;
        L.recs.push_back({dir, k, DS, time_c, tau, c->cs, dir == 0 ? c->cx : c->cy});
        ++L.sweeps;
        int rc = enqueue_sweep(c, dir, tf, nu, slot, prev);
        return rc ? rc : 0;
    };

    bool ended = false;
    for (; k <= N; DS ? k : ++k) {
        if (!skip_frame && time_c >= time_plot[nt_plot] && nt_plot < (N_plot - 1)) {   // plot frame
            const int f = flush();
            if (f < 0) return -2;
            if (f) { ended = true; break; }
            if (nt == 0 && nt < N_T - 1) {   // the copy freezes CV[0]: the CFL bound of its last state
                const int rc = refresh_cfl(c);
                if (rc) return rc;
            }
            c->at = {1, k, DS, nt < (N_T - 1) ? nt + 1 : nt, nt_plot + 1, time_c, tau, half_tau, half_nu, mu};
            const int stop = fn ? fn(user, nt, nt_plot, time_c) : 0;
            c->at.valid = 0;
            if (stop) { status = 1; ended = true; break; }
            nt_plot++;
            if (nt < (N_T - 1)) nt++;
        }
        skip_frame = false;
        if (DS) {                                                             // time step
            if (nt == 0) {                    // no frame copied off yet: CV[0] is the live state
                const int f = flush();
                if (f < 0) return -2;
                if (f) { ended = true; break; }
                const int rc = refresh_cfl(c);
                if (rc) return rc;
            }
            const double h_S_max = c->h_S_max;
            if (isfinite(t_all) || !isfinite(c->cfg[16]) || c->cfg[16] <= 0.0) {
                tau = CFL * h_S_max;
                if (tau < eps) {
                    printf("\nThe length of the time step is so small on [%d, %g, %g] (t_n, time_c, tau)\n", k, time_c, tau);
                    stop_t = true;
                } else if ((time_c + tau) > (t_all - eps)) {
                    tau = t_all - time_c;
                } else if (!isfinite(tau)) {
                    printf("NAN or INFinite error on [%d, %g, %g] (t_n, time_c, tau) - CFL\n", k, time_c, tau);
                    const int f = flush();
                    if (f < 0) return -2;
                    if (!f) status = 1;
                    ended = true;
                    break;
                }
            }
            half_tau = tau * 0.5;
            half_nu = half_tau / h_x;
            mu = tau / h_y;
        }
        if (!valid_bc(bc_x)) {
            printf("No suitable boundary coditions in x direction!\n");
            const int f = flush();
            if (f < 0) return -2;
            if (!f) status = 1;
            ended = true;
            break;
        }
        int rc = sweep(0, half_tau, half_nu);
        if (rc < 0) return rc;
        if (rc > 0) { ended = true; break; }
        if (stop_t) break;
        if (DS) {
            if (!valid_bc(bc_y)) {
                printf("No suitable boundary coditions in y direction!\n");
                const int f = flush();
                if (f < 0) return -2;
                if (!f) status = 1;
                ended = true;
                break;
            }
            rc = sweep(1, tau, mu);
            if (rc < 0) return rc;
            if (rc > 0) { ended = true; break; }
            time_c += tau;
        }
        if (time_c > (t_all - eps) || !isfinite(time_c)) break;
        DS = DS ? 0 : 1;
    }
    if (!ended) {
        const int f = flush();
        if (f < 0) return -2;
        if (stop_t && status == 0) status = 2;
    }
    HCHECK(hipStreamSynchronize(c->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (status != 1) {
        printf("\nTime is up at time step %d.\n", k);
        printf("The cost of CPU time for 2D-GRP Eulerian scheme with dimension splitting for this problem is %g seconds.\n", secs);
    }
    res->k = k;
    res->nt = nt;
    res->status = status;
    res->sweeps = L.sweeps;
    res->time_c = time_c;
    res->tau = tau;
    res->seconds = secs;
    return 0;
}

// ---- the unsplit scheme ----------------------------------------------------------
// One launch per step; status slots come in pairs ([0] x-interfaces and the
// update, [1] y-interfaces).
static constexpr int UNSPLIT_SLOTS = STATUS_SLOTS / 2;

static void fill_unsplit_args(hgrp2d *c, UnsplitArgs &a, double tau, double nu, double mu, int slot, int prev_slot)
{
    const int in = c->cs, out = 1 - c->cs;
    // the planes of each group are consecutive in c->block (hgrp2d_create)
    a.st = c->st[in][0];
    a.st_o = c->st[out][0];
    a.s = c->sx[c->cx][0]; a.s_o = c->sx[1 - c->cx][0];
    a.t = c->sy[c->cy][0]; a.t_o = c->sy[1 - c->cy][0];
    a.plane = (size_t)c->m * c->ld;
    a.bx_lo = c->bnd[0]; a.bx_hi = c->bnd[1]; a.by_lo = c->bnd[2]; a.by_hi = c->bnd[3];
    a.m = c->m; a.n = c->n; a.ld = c->ld;
    a.bc_x = (int)c->cfg[17]; a.bc_y = (int)c->cfg[18];
    a.lim3_x = c->found_x; a.lim3_y = c->found_y;
    a.hh_x = 0.5 * c->cfg[10]; a.hh_y = 0.5 * c->cfg[11];
    a.div_hx = recip_div(c->cfg[10]); a.div_hy = recip_div(c->cfg[11]);
    a.tau = tau; a.nu = nu; a.mu = mu;
    a.eps = c->cfg[4];
    a.alpha = c->cfg[41];
    a.tol = c->newton_tol > 0.0 ? c->newton_tol : a.eps;
    a.conv = c->newton_rule == HGRP_NEWTON_CONVERGE;
    a.g2d = c->g2d;
    a.form = c->uform;
    a.mlp = c->use_mlp ? c->mlp : nullptr;
    a.gas = make_gas(c->cfg[6]);
    a.status = c->status + 2 * slot;
    a.prev = prev_slot >= 0 ? c->status + 2 * prev_slot : nullptr;
}

static int enqueue_unsplit(hgrp2d *c, double tau, double nu, double mu, int slot, int prev_slot)
{
    UnsplitArgs a;
    fill_unsplit_args(c, a, tau, nu, mu, slot, prev_slot);
    HCHECK(launch_unsplit_step(a, c->stream));
    c->cs = 1 - c->cs;
    c->cx = 1 - c->cx;
    c->cy = 1 - c->cy;
    c->found_x = c->found_y = 1;
    return 0;
}

static bool unsplit_supported(hgrp2d *c)
{
    if (c->two_fluid) {
        printf("The unsplit scheme (config[33] = 0) runs one gamma-law gas only.\n");
        return false;
    }
    if (c->m > unsplit_max_rows()) {
        printf("The unsplit scheme takes at most %d cells in x (n_x = %d).\n", unsplit_max_rows(), c->m);
        return false;
    }
    return true;
}

// GRP_solver_2D_EUL_source (src/finite_volume/grp_solver_2D_EUL_source.c:147-300)
static int solve_unsplit(hgrp2d *c, const double *time_plot, int N_plot, int N_T, hgrp2d_frame_fn fn, void *user,
                         hgrp2d_result *res)
{
    if (!unsplit_supported(c)) return ARG_ERROR - 4;
    const double t_all = c->cfg[1], eps = c->cfg[4], CFL = c->cfg[7];
    const int N = (int)c->cfg[5];
    const double h_x = c->cfg[10], h_y = c->cfg[11];
    double tau = c->cfg[16];
    const int bc_x = (int)c->cfg[17], bc_y = (int)c->cfg[18];
    const auto t_start = std::chrono::steady_clock::now();

    int k = 1, nt = 0, nt_plot = 0, status = 0, launched = 0;
    double time_c = 0.0, nu = 0.0, mu = 0.0;
    bool stop_t = false, skip_frame = false, ended = false;
    if (c->resume.valid) {
        const LoopPos &r = c->resume;
        k = r.k; nt = r.nt; nt_plot = r.nt_plot;
        time_c = r.time_c; tau = r.tau; nu = r.half_nu; mu = r.mu;
        skip_frame = true;
        c->resume.valid = 0;
        if (nt > N_T - 1) nt = N_T - 1;
    }
    HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));

    struct Rec { int k; double time_c, tau; int in_s; };
    std::vector<Rec> recs;
    // reads the pending steps' status; at the first failed one, restores the
    // loop state it ran in and applies the reference's rule.  Returns 1 if a
    // failure ended the run, 0 if none, -2 on a HIP error.
    auto flush = [&]() -> int {
        if (recs.empty()) return 0;
        std::vector<SweepStatus> st(2 * recs.size());
        if (hipMemcpyAsync(st.data(), c->status, st.size() * sizeof(SweepStatus), hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return -2;
        for (size_t r = 0; r < recs.size(); ++r) {
            const SweepStatus &x = st[2 * r], &y = st[2 * r + 1];
            if (!x.fail_key && !y.fail_key && !(x.flags & 1u)) continue;
            const int cx = x.fail_key ? (int)((x.fail_key - 1ull) & 3ull) : 0;
            const int cy = y.fail_key ? (int)((y.fail_key - 1ull) & 3ull) : 0;
            const Rec &R = recs[r];
            k = R.k; time_c = R.time_c; tau = R.tau;
            launched += (int)r + 1 - (int)recs.size();
            if (cx == 1 || cy == 1) {                   // flux_generator_{x,y} returned 1: goto return_NULL
                printf("<0.0 or non-finite reconstruction at the %c-interfaces of step %d\n", cx == 1 ? 'x' : 'y', k);
                c->cs = R.in_s;
                status = 1;
            } else {                                    // GRP or positivity failure: update kept, stop
                if (cx == 2) printf("GRP star/derivative error at the x-interfaces of step %d\n", k);
                if (cy == 2) printf("GRP star/derivative error at the y-interfaces of step %d\n", k);
                if (x.flags & 1u) printf("<0.0 error on step %d - Update\n", k);
                c->cs = 1 - R.in_s;
                time_c += R.tau;
                status = 2;
            }
            return 1;
        }
        recs.clear();
        return 0;
    };

    for (; k <= N; ++k) {
        if (!skip_frame && time_c >= time_plot[nt_plot] && nt_plot < (N_plot - 1)) {   // plot frame (:157-176)
            const int f = flush();
            if (f < 0) return -2;
            if (f) { ended = true; break; }
            if (nt == 0 && nt < N_T - 1) {   // the copy freezes CV[0]: the CFL bound of its last state
                const int rc = refresh_cfl(c);
                if (rc) return rc;
            }
            c->at = {1, k, 1, nt < (N_T - 1) ? nt + 1 : nt, nt_plot + 1, time_c, tau, 0.0, nu, mu};
            const int stop = fn ? fn(user, nt, nt_plot, time_c) : 0;
            c->at.valid = 0;
            if (stop) { status = 1; ended = true; break; }
            nt_plot++;
            if (nt < (N_T - 1)) nt++;
        }
        skip_frame = false;
        if (nt == 0) {                     // CV[0] is the live state: its CFL bound moves (:183-191)
            const int f = flush();
            if (f < 0) return -2;
            if (f) { ended = true; break; }
            const int rc = refresh_cfl(c);
            if (rc) return rc;
        }
        if (isfinite(t_all) || !isfinite(c->cfg[16]) || c->cfg[16] <= 0.0) {          // (:193-207)
            tau = CFL * c->h_S_max;
            if (tau < eps) {
                printf("\nThe length of the time step is so small on [%d, %g, %g] (t_n, time_c, tau)\n", k, time_c, tau);
                stop_t = true;
            } else if ((time_c + tau) > (t_all - eps)) {
                tau = t_all - time_c;
            } else if (!isfinite(tau)) {
                printf("NAN or INFinite error on [%d, %g, %g] (t_n, time_c, tau) - CFL\n", k, time_c, tau);
                const int f = flush();
                if (f < 0) return -2;
                if (!f) status = 1;
                ended = true;
                break;
            }
        }
        nu = tau / h_x;
        mu = tau / h_y;
        if (!valid_bc(bc_x) || !valid_bc(bc_y)) {
            printf("No suitable boundary coditions in %c direction!\n", valid_bc(bc_x) ? 'y' : 'x');
            const int f = flush();
            if (f < 0) return -2;
            if (!f) status = 1;
            ended = true;
            break;
        }
        if ((int)recs.size() + 1 > UNSPLIT_SLOTS) {
            const int f = flush();
            if (f < 0) return -2;
            if (f) { ended = true; break; }
        }
        const int slot = (int)recs.size();
        recs.push_back({k, time_c, tau, c->cs});
        ++launched;
        const int rc = enqueue_unsplit(c, tau, nu, mu, slot, slot - 1);
        if (rc) return rc;
        time_c += tau;
        if (stop_t || time_c > (t_all - eps) || !isfinite(time_c)) break;
    }
    if (!ended) {
        const int f = flush();
        if (f < 0) return -2;
        if (f == 0 && stop_t) status = 2;
    }
    HCHECK(hipStreamSynchronize(c->stream));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (status != 1) {
        printf("\nTime is up at time step %d.\n", k);
        printf("The cost of CPU time for genuinely 2D-GRP Eulerian scheme without dimension splitting for this problem is %g seconds.\n", secs);
    }
    res->k = k;
    res->nt = nt;
    res->status = status;
    res->sweeps = launched;
    res->time_c = time_c;
    res->tau = tau;
    res->seconds = secs;
    return 0;
}

// The unsplit step's autotuner: times the tile and the strip form over
// `steps` steps from the context's current state, two trials each (device
// time of each launch from HIP events), keeps the faster and puts the state
// back, as hgrp2d_tune does for the split sweeps.
static int tune_unsplit(hgrp2d *c, int steps, int *form_out, double *ms_out)
{
    if (!valid_bc((int)c->cfg[17]) || !valid_bc((int)c->cfg[18])) return ARG_ERROR - 4;
    if (!unsplit_supported(c)) return ARG_ERROR - 4;
    const size_t cells = (size_t)c->m * c->ld;
    const size_t nb = (size_t)(2 * 5 + 2 * 4 + 2 * 4) * cells;     // state and slope planes
    double *save = nullptr;
    HCHECK(hipMalloc((void **)&save, nb * sizeof(double)));
    if (hipMemcpyAsync(save, c->block, nb * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
        (void)hipFree(save);
        return ARG_ERROR - 5;
    }
    const int cs = c->cs, cx = c->cx, cy = c->cy, fx = c->found_x, fy = c->found_y, form0 = c->uform;
    auto put_back = [&]() -> int {
        HCHECK(hipMemcpyAsync(c->block, save, nb * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        c->cs = cs; c->cx = cx; c->cy = cy; c->found_x = fx; c->found_y = fy;
        return 0;
    };
    const double tau = c->cfg[7] * c->h_S_max;
    const double nu = tau / c->cfg[10], mu = tau / c->cfg[11];
    std::vector<hipEvent_t> evs((size_t)2 * steps);
    for (auto &e : evs)
        if (hipEventCreate(&e) != hipSuccess) { (void)hipFree(save); return ARG_ERROR - 5; }
    auto trial = [&](double *ms) -> int {
        if (int r = put_back()) return r;
        HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));
        for (int s = 0; s < steps; ++s) {
            HCHECK(hipEventRecord(evs[(size_t)2 * s], c->stream));
            if (int r = enqueue_unsplit(c, tau, nu, mu, s % UNSPLIT_SLOTS, -1)) return r;
            HCHECK(hipEventRecord(evs[(size_t)2 * s + 1], c->stream));
        }
        HCHECK(hipEventSynchronize(evs[(size_t)2 * steps - 1]));
        double sum = 0.0;
        for (int s = 0; s < steps; ++s) {
            float t = 0.f;
            HCHECK(hipEventElapsedTime(&t, evs[(size_t)2 * s], evs[(size_t)2 * s + 1]));
            sum += t;
        }
        *ms = sum / steps;
        return 0;
    };
    int rc = 0, pick = HGRP_UNSPLIT_TILE;
    double best = 0.0;
    for (int form = HGRP_UNSPLIT_TILE; form <= HGRP_UNSPLIT_STRIP && !rc; ++form) {
        c->uform = form;
        double t = 1e300;
        for (int k = 0; k < 2 && !rc; ++k) {
            double ms = 0.0;
            rc = trial(&ms);
            if (!rc && ms < t) t = ms;
        }
        if (!rc && (form == HGRP_UNSPLIT_TILE || t < best)) { best = t; pick = form; }
    }
    c->uform = rc ? form0 : pick;
    const int rb = put_back();
    if (!rc) rc = rb;
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = -(int)e - 1;
    }
    for (auto &e : evs) (void)hipEventDestroy(e);
    (void)hipFree(save);
    if (rc) return rc;
    if (form_out) *form_out = c->uform;
    if (ms_out) ms_out[0] = best;
    return 0;
}

extern "C" int hgrp2d_bench_unsplit(hgrp2d *c, int steps, int sample_every, double *ms_total, double *ms_step)
{
    if (!c || !c->uploaded || steps < 1) return ARG_ERROR - 1;
    if (!valid_bc((int)c->cfg[17]) || !valid_bc((int)c->cfg[18])) return ARG_ERROR - 4;
    if (!unsplit_supported(c)) return ARG_ERROR - 4;
    const double tau = c->cfg[7] * c->h_S_max;
    const double nu = tau / c->cfg[10], mu = tau / c->cfg[11];
    HCHECK(hipMemsetAsync(c->status, 0, STATUS_SLOTS * sizeof(SweepStatus), c->stream));
    const int every = sample_every > 0 ? sample_every : steps + 1;
    const int ns = sample_every > 0 ? (steps + every - 1) / every : 0;
    std::vector<hipEvent_t> evs((size_t)2 * ns);
    for (auto &e : evs) HCHECK(hipEventCreate(&e));
    HCHECK(hipEventRecord(c->ev[0], c->stream));
    for (int s = 0; s < steps; ++s) {
        const bool sampled = ns > 0 && s % every == 0;
        if (sampled) HCHECK(hipEventRecord(evs[(size_t)2 * (s / every)], c->stream));
        // status slots are not chained here: the benchmark must run every step
        const int rc = enqueue_unsplit(c, tau, nu, mu, s % UNSPLIT_SLOTS, -1);
        if (rc) return rc;
        if (sampled) HCHECK(hipEventRecord(evs[(size_t)2 * (s / every) + 1], c->stream));
    }
    HCHECK(hipEventRecord(c->ev[1], c->stream));
    HCHECK(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (ms_total) *ms_total = ms;
    double sum = 0.0;
    for (int k = 0; k < ns; ++k) {
        float t = 0.f;
        HCHECK(hipEventElapsedTime(&t, evs[(size_t)2 * k], evs[(size_t)2 * k + 1]));
        sum += t;
    }
    if (ms_step) *ms_step = ns ? sum / ns : NAN;
    for (auto &e : evs) (void)hipEventDestroy(e);
    return 0;
}
