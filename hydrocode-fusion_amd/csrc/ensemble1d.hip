// ensemble1d.hip -- batched 1D Eulerian GRP: many independent problems of
// uneven sizes advanced together, each with its own time step.
//
// One step of GRP_solver_EUL_source (src/finite_volume/grp_solver_EUL_source.c:27-334)
// for every active instance is three launches:
//   iface_kernel   ghosts + minmod slopes of the two cells, reconstruction,
//                  checks, linear_GRP_solver_Edir; interface values and time
//                  derivatives to HBM; per-instance min of h/S (:133-214)
//   update_kernel  per cell: tau from CFL * h/S_max with the reference's clip
//                  and stop rules (:218-233, `step_tau`, the same in every wave
//                  of the instance), the two interface fluxes at t + tau/2, the
//                  conservative update, positivity check, new slopes (:236-277)
//   close_kernel   per instance: the same tau, time, step counter, end and stop
//                  rules (:279-283)
// Waves of one launch share nothing but atomics into the instance's flags and
// h/S minimum, which only the next launch reads.
// The two wide kernels map waves to (instance, first index) through per-wave
// tables built once per active set: every instance starts on a wave boundary,
// so a wave works on one instance only, reads that instance's fields as
// uniform (scalar) loads and takes its h/S minimum with one atomic, and a
// launch covers the cells of the instances still running (plus the unused
// lanes of each instance's last wave).  Steps are captured in hipGraphs of
// STEPS_PER_GRAPH steps; the host
// reads the instances' done flags between graphs and rebuilds the tables (and
// the graph) when enough work has finished.  Block sizes of the two wide
// kernels are picked by a small autotuner that times candidates on the real
// data (hgrp1d_tune).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/hydrocode_grp.h"
#include "grp1d_device.hpp"

using namespace hgrp;

#define ECHECK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "hydrogrp: %s failed: %s (%s:%d)\n", #x,  \
                    hipGetErrorString(e_), __FILE__, __LINE__);       \
            return -(int)e_ - 1;                                      \
        }                                                             \
    } while (0)

namespace {

constexpr int STEPS_PER_GRAPH = 32;
constexpr unsigned FLAG_ABORT = 1u, FLAG_STOP = 2u;

// per-instance state on the device
struct Inst {
    int coff, ioff, m, bc;        // first cell, first interface, cells, boundary condition
    int k, done, status, found;   // step counter (the reference's k), finished, 0/1/2, limiter switch
    unsigned flags;               // FLAG_* raised by the interfaces of the current step
    unsigned uflags;              // FLAG_STOP raised by the update of the current step
    int paused;                   // time_c reached t_pause at the start of a step: waits for the host
    double h, time_c, tau;
    double t_pause;               // plot time the host waits for (INFINITY: none)
    unsigned long long hsm_bits;  // min over interfaces of h/S, as bits (positive doubles order like them)
    double b0[6];                 // BC -1 ghosts from the initial data: rho u p left, rho u p right
};

struct Params {
    double t_all, eps, gamma, cfl, tau_fixed, alpha;
    int n_max;
    Gas gas;
};

__device__ __forceinline__ void flag(Inst *I, unsigned f) { atomicOr(&I->flags, f); }

// values (rho, u, p) of cell c of instance I, ghosts at c = -1 and c = m
__device__ __forceinline__ void cell_value(const Inst &I, const double *rho, const double *u, const double *p, int c,
                                           double w[3])
{
    const int m = I.m;
    if (c >= 0 && c < m) {
        const size_t o = (size_t)I.coff + c;
        w[0] = rho[o]; w[1] = u[o]; w[2] = p[o];
        return;
    }
    const bool lo = c < 0;
    if (I.bc == -1) {
        const int b = lo ? 0 : 3;
        w[0] = I.b0[b]; w[1] = I.b0[b + 1]; w[2] = I.b0[b + 2];
        return;
    }
    int src;
    if (I.bc == -7) src = lo ? m - 1 : 0;
    else src = lo ? 0 : m - 1;
    const size_t o = (size_t)I.coff + src;
    w[0] = rho[o]; w[1] = u[o]; w[2] = p[o];
    if (I.bc == -2 || (I.bc == -24 && lo)) w[1] = -w[1];
}

// limited slopes (rho, u, p) of cell c from the values of cells c - 1, c and
// c + 1 (bound_cond_slope_limiter.c:83-92 with minmod_limiter)
__device__ __forceinline__ void limit3(const Inst &I, const Params &P, const double *const s_prev[3], int c,
                                       const double vm[3], const double v0[3], const double vp[3], double s[3])
{
    const double h = I.h;
    for (int q = 0; q < 3; ++q) {
        const double s_L = (v0[q] - vm[q]) / h;
        const double s_R = (vp[q] - v0[q]) / h;
        s[q] = I.found ? minmod3(P.alpha * s_L, P.alpha * s_R, s_prev[q][(size_t)I.coff + c]) : minmod2(s_L, s_R);
    }
}

// the same, loading the three values
__device__ __forceinline__ void limited(const Inst &I, const Params &P, const double *rho, const double *u,
                                        const double *p, const double *const s_prev[3], int c, double s[3])
{
    double vm[3], v0[3], vp[3];
    cell_value(I, rho, u, p, c - 1, vm);
    cell_value(I, rho, u, p, c, v0);
    cell_value(I, rho, u, p, c + 1, vp);
    limit3(I, P, s_prev, c, vm, v0, vp, s);
}

// ghost slopes (bound_cond_slope_limiter.c:94-109): -2 u only, -7 all, -24 left u only
__device__ __forceinline__ void ghost_slope(const Inst &I, const Params &P, const double *rho, const double *u,
                                            const double *p, const double *const s_prev[3], bool lo, double s[3])
{
    s[0] = s[1] = s[2] = 0.0;
    double t[3];
    if (I.bc == -2) {
        limited(I, P, rho, u, p, s_prev, lo ? 0 : I.m - 1, t);
        s[1] = t[1];
    } else if (I.bc == -7) {
        limited(I, P, rho, u, p, s_prev, lo ? I.m - 1 : 0, s);
    } else if (I.bc == -24 && lo) {
        limited(I, P, rho, u, p, s_prev, 0, t);
        s[1] = t[1];
    }
}

struct Tables {
    const int2 *w_if, *w_c;           // per wave of iface_kernel / update_kernel: {instance id, first index}
    const int *act;                   // active instance ids
};

// The wave's index in a wide launch, as a uniform value.
template <int BS>
__device__ __forceinline__ int wave_index()
{
    return __builtin_amdgcn_readfirstlane((int)((blockIdx.x * BS + threadIdx.x) >> 6));
}

// four waves per SIMD: 128 VGPRs and a 4-register spill (unbounded it takes 130, three waves);
// measured 5 % faster on the 512-instance ensemble
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4))) iface_kernel(Tables T, int n_w, Inst *inst, Params P, const double *rho,
                                                   const double *u, const double *p, const double *s_rho,
                                                   const double *s_u, const double *s_p, double *mid, double *dire,
                                                   size_t n_if_all)
{
    const int wv = wave_index<BS>();
    if (wv >= n_w) return;
    const int2 we = T.w_if[wv];
    Inst *Ip = &inst[we.x];
    if (Ip->done || Ip->paused) return;
    const Inst &I = *Ip;
    const int m = I.m;
    const double *const sp[3] = {s_rho, s_u, s_p};
    const int t = (int)(threadIdx.x & 63);
    const int j = we.y + t;                         // interface j: between cells j - 1 and j
    // Each lane holds the left cell of its interface, c = j - 1 (-1 and m are
    // the ghosts): its value, then its slope, limited with the neighbours'
    // values from the previous and next lanes (the wave's end lanes load
    // theirs).  The right cell's value and slope come from the next lane; the
    // last lane forms its right cell's slope itself.
    const int c = j - 1;
    const bool interior = c >= 0 && c < m;
    double wl[3] = {1.0, 0.0, 1.0}, wr[3], vm[3], sl[3] = {0.0, 0.0, 0.0}, sr[3];
    if (c <= m) cell_value(I, rho, u, p, c, wl);
    for (int q = 0; q < 3; ++q) { vm[q] = __shfl_up(wl[q], 1); wr[q] = __shfl_down(wl[q], 1); }
    if (t == 0 && interior) cell_value(I, rho, u, p, c - 1, vm);
    if (t == 63 && interior) cell_value(I, rho, u, p, c + 1, wr);
    if (interior) limit3(I, P, sp, c, vm, wl, wr, sl);
    else if (c == -1) ghost_slope(I, P, rho, u, p, sp, true, sl);
    else if (c == m) ghost_slope(I, P, rho, u, p, sp, false, sl);
    for (int q = 0; q < 3; ++q) sr[q] = __shfl_down(sl[q], 1);
    if (t == 63 && j < m) limited(I, P, rho, u, p, sp, j, sr);
    else if (t == 63 && j == m) ghost_slope(I, P, rho, u, p, sp, false, sr);
    double hs = INFINITY;
    if (j <= m) {
        const double h = I.h, gamma = P.gamma;
        Side1 L, R;
        L.rho = wl[0] + 0.5 * h * sl[0]; L.u = wl[1] + 0.5 * h * sl[1]; L.p = wl[2] + 0.5 * h * sl[2];
        if (j < m) {
            R.rho = wr[0] - 0.5 * h * sr[0]; R.u = wr[1] - 0.5 * h * sr[1]; R.p = wr[2] - 0.5 * h * sr[2];
        } else {   // the reference reconstructs the right ghost with +h/2 (grp_solver_EUL_source.c:154-158)
            R.rho = wr[0] + 0.5 * h * sr[0]; R.u = wr[1] + 0.5 * h * sr[1]; R.p = wr[2] + 0.5 * h * sr[2];
        }
        L.s_rho = sl[0]; L.s_u = sl[1]; L.s_p = sl[2];
        R.s_rho = sr[0]; R.s_u = sr[1]; R.s_p = sr[2];
        const double c_L = sqrt(gamma * L.p / L.rho);
        const double c_R = sqrt(gamma * R.p / R.rho);
        hs = fmin(hs, h / (fabs(L.u) + fabs(c_L)));
        hs = fmin(hs, h / (fabs(R.u) + fabs(c_R)));
        const double eps = P.eps;
        if (L.p < eps || R.p < eps || L.rho < eps || R.rho < eps || !isfinite(L.s_p) || !isfinite(R.s_p) ||
            !isfinite(L.s_u) || !isfinite(R.s_u) || !isfinite(L.s_rho) || !isfinite(R.s_rho)) {
            flag(Ip, FLAG_ABORT);
        } else {
            double D[3], U[3];
            grp_edir(D, U, L, R, P.gas, eps);
            if (U[2] < eps || U[0] < eps || !isfinite(U[1]) || !isfinite(U[2]) || !isfinite(U[0]) ||
                !isfinite(D[1]) || !isfinite(D[2]) || !isfinite(D[0]))
                flag(Ip, FLAG_STOP);
            const size_t g = (size_t)I.ioff + j;
            for (int q = 0; q < 3; ++q) { mid[q * n_if_all + g] = U[q]; dire[q * n_if_all + g] = D[q]; }
        }
    }
    // per-instance min of h/S: one atomic per wave
    double wmin = hs;
    for (int off = 32; off > 0; off >>= 1) wmin = fmin(wmin, __shfl_xor(wmin, off));
    if ((threadIdx.x & 63) == 0 && wmin < INFINITY)
        atomicMin(&Ip->hsm_bits, (unsigned long long)__double_as_longlong(wmin));
}

// tau of instance I for the step (grp_solver_EUL_source.c:218-233) from the
// h/S minimum its interfaces left: `stop` where tau < eps (the step is taken,
// then the run stops), `bad` where tau is not finite (no update, error).
__device__ __forceinline__ double step_tau(const Inst &I, const Params &P, bool &stop, bool &bad)
{
    stop = bad = false;
    double tau = I.tau;
    if (isfinite(P.t_all) || !isfinite(P.tau_fixed) || P.tau_fixed <= 0.0) {
        const double h_S_max = __longlong_as_double((long long)I.hsm_bits);
        tau = P.cfl * h_S_max;
        if (tau < P.eps) stop = true;
        else if ((I.time_c + tau) > (P.t_all - P.eps)) tau = P.t_all - I.time_c;
        else if (!isfinite(tau)) bad = true;
    }
    return tau;
}

__device__ __forceinline__ void iface_flux(const double *mid, const double *dire, size_t n_if_all, size_t g,
                                           double tau, double gamma, double F[3], double In[3])
{
    double R = mid[g], U = mid[n_if_all + g], Pp = mid[2 * n_if_all + g];
    const double Rt = dire[g], Ut = dire[n_if_all + g], Pt = dire[2 * n_if_all + g];
    R += 0.5 * tau * Rt;
    U += 0.5 * tau * Ut;
    Pp += 0.5 * tau * Pt;
    F[0] = R * U;
    F[1] = F[0] * U + Pp;
    F[2] = (gamma / (gamma - 1.0)) * Pp + 0.5 * F[0] * U;
    F[2] = F[2] * U;
    In[0] = R + 0.5 * tau * Rt;
    In[1] = U + 0.5 * tau * Ut;
    In[2] = Pp + 0.5 * tau * Pt;
}

template <int BS>
__global__ void __launch_bounds__(BS) update_kernel(Tables T, int n_w, Inst *inst, Params P, double *rho, double *u,
                                                    double *p, double *e, double *s_rho, double *s_u, double *s_p,
                                                    const double *mid, const double *dire, size_t n_if_all)
{
    const int wv = wave_index<BS>();
    if (wv >= n_w) return;
    const int2 we = T.w_c[wv];
    Inst &I = inst[we.x];
    if (I.done || I.paused) return;
    const int j = we.y + (int)(threadIdx.x & 63);   // cell j
    if (j >= I.m) return;
    if (I.flags & FLAG_ABORT) return;   // reconstruction error: no update (close_kernel ends the run)
    bool stop, bad;
    const double tau = step_tau(I, P, stop, bad);
    if (bad) return;
    const double h = I.h, gamma = P.gamma;
    const double nu = tau / h;
    const size_t g = (size_t)I.ioff + j;
    double Fl[3], Il[3], Fr[3], Ir[3];
    iface_flux(mid, dire, n_if_all, g, tau, gamma, Fl, Il);
    iface_flux(mid, dire, n_if_all, g + 1, tau, gamma, Fr, Ir);
    const size_t o = (size_t)I.coff + j;
    const double r0 = rho[o], u0 = u[o], e0 = e[o];
    const double Mom = r0 * u0 - nu * (Fr[1] - Fl[1]);
    const double Ene = r0 * e0 - nu * (Fr[2] - Fl[2]);
    const double r1 = r0 - nu * (Fr[0] - Fl[0]);
    const double u1 = Mom / r1;
    const double e1 = Ene / r1;
    const double p1 = (Ene - 0.5 * Mom * u1) * (gamma - 1.0);
    rho[o] = r1; u[o] = u1; e[o] = e1; p[o] = p1;
    if (p1 < P.eps || r1 < P.eps) atomicOr(&I.uflags, FLAG_STOP);
    s_u[o] = (Ir[1] - Il[1]) / h;
    s_p[o] = (Ir[2] - Il[2]) / h;
    s_rho[o] = (Ir[0] - Il[0]) / h;
}

// end of a step (grp_solver_EUL_source.c:218-233, 279-283): the step's tau
// again (update_kernel used the same), time, stop and end rules, step counter;
// clears the per-step accumulators.  Then the plot-time test the
// reference makes at the start of the next step (:116-125): an instance whose
// time has reached its t_pause waits for the host, which copies the frame.
__global__ void close_kernel(const int *act, int n_act, Inst *inst, Params P)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_act) return;
    Inst &I = inst[act[s]];
    if (I.done || I.paused) return;
    if (I.flags & FLAG_ABORT) {            // reconstruction error: no update, stop here
        I.done = 1;
        I.status = 1;
        return;
    }
    bool stop, bad;
    const double tau = step_tau(I, P, stop, bad);
    if (bad) {
        I.done = 1;
        I.status = 1;
        return;
    }
    I.tau = tau;
    I.found = 1;
    I.time_c += tau;
    const unsigned f = I.flags | I.uflags | (stop ? FLAG_STOP : 0u);
    I.flags = I.uflags = 0u;
    I.hsm_bits = 0x7ff0000000000000ull;
    if ((f & FLAG_STOP) || I.time_c > (P.t_all - P.eps) || !isfinite(I.time_c)) {
        I.done = 1;
        I.status = (f & FLAG_STOP) ? 2 : 0;
        return;
    }
    if (++I.k > P.n_max) {
        I.done = 1;
        return;
    }
    if (I.time_c >= I.t_pause) I.paused = 1;
}

constexpr int BS_CHOICES[4] = {64, 128, 256, 512};

template <int BS>
void launch_iface(dim3 g, hipStream_t st, const Tables &T, int n, Inst *inst, const Params &P, double **f,
                  double *mid, double *dire, size_t nia)
{
    hipLaunchKernelGGL(iface_kernel<BS>, g, dim3(BS), 0, st, T, n, inst, P, f[0], f[1], f[2], f[4], f[5], f[6], mid,
                       dire, nia);
}
template <int BS>
void launch_update(dim3 g, hipStream_t st, const Tables &T, int n, Inst *inst, const Params &P, double **f,
                   const double *mid, const double *dire, size_t nia)
{
    hipLaunchKernelGGL(update_kernel<BS>, g, dim3(BS), 0, st, T, n, inst, P, f[0], f[1], f[2], f[3], f[4], f[5], f[6],
                       mid, dire, nia);
}

}  // namespace

struct hgrp1d {
    int n_inst;
    size_t n_cells, n_if;
    std::vector<int> m, coff, ioff;
    std::vector<double> h;
    Params P;
    hipStream_t stream = nullptr;
    double *block = nullptr;           // rho u p e s_rho s_u s_p (n_cells each), mid, dire (3 n_if each), initial copy
    double *f[7];
    double *mid, *dire, *init;         // init: rho u p e of the upload (for restarts)
    Inst *inst = nullptr;
    int *tab = nullptr;                // w_if[max_wi], w_c[max_wc] (int2 each), act[n_inst]
    size_t max_wi = 0, max_wc = 0;     // waves of the two wide launches with every instance active
    Tables T{};
    int n_act = 0, n_wi = 0, n_wc = 0, act_c = 0;
    int bs_if = 128, bs_up = 256;      // block sizes (autotuned)
    bool uploaded = false;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<Inst> host_inst;
};

static int bs_index(int bs)
{
    for (int q = 0; q < 4; ++q)
        if (BS_CHOICES[q] == bs) return q;
    return -1;
}

static void enqueue_step(hgrp1d *c)
{
    const dim3 gi((unsigned)(((size_t)c->n_wi * 64 + c->bs_if - 1) / c->bs_if)),
        gu((unsigned)(((size_t)c->n_wc * 64 + c->bs_up - 1) / c->bs_up));
    const dim3 ga((c->n_act + 127) / 128);
    switch (c->bs_if) {
    case 64: launch_iface<64>(gi, c->stream, c->T, c->n_wi, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    case 256: launch_iface<256>(gi, c->stream, c->T, c->n_wi, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    case 512: launch_iface<512>(gi, c->stream, c->T, c->n_wi, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    default: launch_iface<128>(gi, c->stream, c->T, c->n_wi, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    }
    switch (c->bs_up) {
    case 64: launch_update<64>(gu, c->stream, c->T, c->n_wc, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    case 128: launch_update<128>(gu, c->stream, c->T, c->n_wc, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    case 512: launch_update<512>(gu, c->stream, c->T, c->n_wc, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    default: launch_update<256>(gu, c->stream, c->T, c->n_wc, c->inst, c->P, c->f, c->mid, c->dire, c->n_if); break;
    }
    hipLaunchKernelGGL(close_kernel, ga, dim3(128), 0, c->stream, c->T.act, c->n_act, c->inst, c->P);
}

static void drop_graph(hgrp1d *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// rebuilds the active set from host_inst (done and paused flags) and the lookup tables
static int rebuild_tables(hgrp1d *c)
{
    std::vector<int> act;
    std::vector<int2> w_if, w_c;   // per wave: {instance, first interface / cell}
    int act_c = 0;
    for (int i = 0; i < c->n_inst; ++i)
        if (!c->host_inst[i].done && !c->host_inst[i].paused) {
            act.push_back(i);
            act_c += c->m[i];
            for (int j = 0; j <= c->m[i]; j += 64) w_if.push_back(make_int2(i, j));   // interfaces 0 .. m
            for (int j = 0; j < c->m[i]; j += 64) w_c.push_back(make_int2(i, j));     // cells 0 .. m - 1
        }
    c->n_act = (int)act.size();
    c->n_wi = (int)w_if.size();
    c->n_wc = (int)w_c.size();
    c->act_c = act_c;
    drop_graph(c);
    if (!c->n_act) return 0;
    int2 *dw_if = (int2 *)c->tab, *dw_c = dw_if + c->max_wi;
    int *dact = (int *)(dw_c + c->max_wc);
    ECHECK(hipMemcpyAsync(dw_if, w_if.data(), w_if.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    ECHECK(hipMemcpyAsync(dw_c, w_c.data(), w_c.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    ECHECK(hipMemcpyAsync(dact, act.data(), act.size() * 4, hipMemcpyHostToDevice, c->stream));
    c->T.w_if = dw_if;
    c->T.w_c = dw_c;
    c->T.act = dact;
    ECHECK(hipStreamSynchronize(c->stream));   // the host vectors go out of scope
    return 0;
}

static int reset_instances(hgrp1d *c)
{
    for (int i = 0; i < c->n_inst; ++i) {
        Inst &I = c->host_inst[i];
        I.k = 1; I.done = 0; I.status = 0; I.found = 0; I.flags = I.uflags = 0u; I.paused = 0;
        I.time_c = 0.0;
        I.t_pause = INFINITY;
        I.tau = c->P.tau_fixed;
        I.hsm_bits = 0x7ff0000000000000ull;
    }
    ECHECK(hipMemcpyAsync(c->inst, c->host_inst.data(), c->n_inst * sizeof(Inst), hipMemcpyHostToDevice, c->stream));
    // state from the upload copy, slopes zero
    for (int q = 0; q < 4; ++q)
        ECHECK(hipMemcpyAsync(c->f[q], c->init + q * c->n_cells, c->n_cells * 8, hipMemcpyDeviceToDevice, c->stream));
    for (int q = 4; q < 7; ++q) ECHECK(hipMemsetAsync(c->f[q], 0, c->n_cells * 8, c->stream));
    ECHECK(hipStreamSynchronize(c->stream));
    return rebuild_tables(c);
}

extern "C" int hgrp1d_create(hgrp1d **out, int n_inst, const int *cells, const double *h, const double *config)
{
    if (!out || n_inst < 1 || !cells || !h || !config) return -1001;
    *out = nullptr;
    const int bc = (int)config[17];
    if (bc != -1 && bc != -2 && bc != -4 && bc != -7 && bc != -24) {
        printf("No suitable boundary coditions in x direction!\n");
        return -1004;
    }
    hgrp1d *c = new hgrp1d();
    c->n_inst = n_inst;
    c->m.assign(cells, cells + n_inst);
    c->h.assign(h, h + n_inst);
    size_t nc = 0, ni = 0;
    for (int i = 0; i < n_inst; ++i) {
        if (cells[i] < 1 || !(h[i] > 0.0)) { delete c; return -1001; }
        c->coff.push_back((int)nc);
        c->ioff.push_back((int)ni);
        nc += (size_t)cells[i];
        ni += (size_t)cells[i] + 1;
        c->max_wi += ((size_t)cells[i] + 64) / 64;   // ceil((m + 1) / 64) waves of interfaces
        c->max_wc += ((size_t)cells[i] + 63) / 64;
    }
    if (ni > 0x7fffffff || c->max_wi > 0x7fffffff / 64) { delete c; return -1001; }   // thread and wave indices fit an int
    c->n_cells = nc;
    c->n_if = ni;
    c->P.t_all = config[1];
    c->P.eps = config[4];
    c->P.n_max = (int)config[5];
    c->P.gamma = config[6];
    c->P.cfl = config[7];
    c->P.tau_fixed = config[16];
    c->P.alpha = config[41];
    c->P.gas = make_gas(config[6]);
    c->host_inst.resize(n_inst);
    for (int i = 0; i < n_inst; ++i) {
        Inst &I = c->host_inst[i];
        memset(&I, 0, sizeof I);
        I.coff = c->coff[i]; I.ioff = c->ioff[i]; I.m = cells[i]; I.bc = bc; I.h = h[i];
    }
    const size_t nb = 7 * nc + 6 * ni + 4 * nc;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&c->block, nb * 8);
    if (e == hipSuccess) e = hipMalloc((void **)&c->inst, n_inst * sizeof(Inst));
    if (e == hipSuccess) e = hipMalloc((void **)&c->tab, (c->max_wi + c->max_wc) * sizeof(int2) + (size_t)n_inst * 4);
    if (e != hipSuccess) {
        fprintf(stderr, "hydrogrp: 1D ensemble allocation failed: %s\n", hipGetErrorString(e));
        hgrp1d_destroy(c);
        return -(int)e - 1;
    }
    for (int q = 0; q < 7; ++q) c->f[q] = c->block + q * nc;
    c->mid = c->block + 7 * nc;
    c->dire = c->mid + 3 * ni;
    c->init = c->dire + 3 * ni;
    *out = c;
    return 0;
}

extern "C" int hgrp1d_destroy(hgrp1d *c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    if (c->block) (void)hipFree(c->block);
    if (c->inst) (void)hipFree(c->inst);
    if (c->tab) (void)hipFree(c->tab);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

extern "C" int hgrp1d_upload(hgrp1d *c, const double *rho, const double *u, const double *p, const double *e)
{
    if (!c || !rho || !u || !p) return -1001;
    const size_t nc = c->n_cells;
    std::vector<double> E(nc);
    const double g = c->P.gamma;
    for (size_t q = 0; q < nc; ++q)   // as src/hydrocode_1D/hydrocode.c:235 forms it
        E[q] = e ? e[q] : 0.5 * u[q] * u[q] + p[q] / (g - 1.0) / rho[q];
    ECHECK(hipMemcpyAsync(c->init, rho, nc * 8, hipMemcpyHostToDevice, c->stream));
    ECHECK(hipMemcpyAsync(c->init + nc, u, nc * 8, hipMemcpyHostToDevice, c->stream));
    ECHECK(hipMemcpyAsync(c->init + 2 * nc, p, nc * 8, hipMemcpyHostToDevice, c->stream));
    ECHECK(hipMemcpyAsync(c->init + 3 * nc, E.data(), nc * 8, hipMemcpyHostToDevice, c->stream));
    for (int i = 0; i < c->n_inst; ++i) {
        Inst &I = c->host_inst[i];
        const size_t a = (size_t)c->coff[i], b = a + (size_t)c->m[i] - 1;
        I.b0[0] = rho[a]; I.b0[1] = u[a]; I.b0[2] = p[a];
        I.b0[3] = rho[b]; I.b0[4] = u[b]; I.b0[5] = p[b];
    }
    ECHECK(hipStreamSynchronize(c->stream));
    c->uploaded = true;
    return reset_instances(c);
}

extern "C" int hgrp1d_download(hgrp1d *c, double *rho, double *u, double *p, double *e)
{
    if (!c || !c->uploaded) return -1001;
    double *dst[4] = {rho, u, p, e};
    for (int q = 0; q < 4; ++q)
        if (dst[q]) ECHECK(hipMemcpyAsync(dst[q], c->f[q], c->n_cells * 8, hipMemcpyDeviceToHost, c->stream));
    ECHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// Captures STEPS_PER_GRAPH steps over the current active set into a graph
// (once per set: rebuild_tables drops it).
static int ensure_graph(hgrp1d *c)
{
    if (c->gexec) return 0;
    ECHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < STEPS_PER_GRAPH; ++s) enqueue_step(c);
    ECHECK(hipStreamEndCapture(c->stream, &c->graph));
    ECHECK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    return 0;
}

// Runs every instance to its end (or max_steps steps).  use_graph 0: plain
// launches (for checks).  The first graph is captured before the timed region.
static int run_all(hgrp1d *c, int use_graph, int max_steps, double *ms_out)
{
    if (use_graph && c->n_act > 0 && (max_steps <= 0 || max_steps >= STEPS_PER_GRAPH)) {
        const int rc = ensure_graph(c);
        if (rc) return rc;
    }
    hipEvent_t e0, e1;
    ECHECK(hipEventCreate(&e0));
    ECHECK(hipEventCreate(&e1));
    ECHECK(hipEventRecord(e0, c->stream));
    int steps = 0, since_rebuild_act = c->act_c;
    while (c->n_act > 0 && (max_steps <= 0 || steps < max_steps)) {
        int chunk = STEPS_PER_GRAPH;
        if (max_steps > 0) chunk = std::min(chunk, max_steps - steps);
        if (use_graph && chunk == STEPS_PER_GRAPH) {
            const int rc = ensure_graph(c);
            if (rc) return rc;
            ECHECK(hipGraphLaunch(c->gexec, c->stream));
        } else {
            for (int s = 0; s < chunk; ++s) enqueue_step(c);
            ECHECK(hipGetLastError());
        }
        steps += chunk;
        ECHECK(hipMemcpyAsync(c->host_inst.data(), c->inst, c->n_inst * sizeof(Inst), hipMemcpyDeviceToHost,
                              c->stream));
        ECHECK(hipStreamSynchronize(c->stream));
        long live_c = 0;
        int live = 0;
        for (int i = 0; i < c->n_inst; ++i)
            if (!c->host_inst[i].done && !c->host_inst[i].paused) { live_c += c->m[i]; ++live; }
        // shrink the launches once a tenth of the covered cells has finished (or everything has)
        if (live == 0 || live_c < 0.9 * since_rebuild_act) {
            int rc = rebuild_tables(c);
            if (rc) return rc;
            since_rebuild_act = c->act_c;
        }
    }
    ECHECK(hipEventRecord(e1, c->stream));
    ECHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    ECHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_out) *ms_out = ms;
    return 0;
}

// Sets every instance's pause time (NULL: none), clears the paused flags and
// rebuilds the active set.
static int arm(hgrp1d *c, const double *t_pause)
{
    ECHECK(hipMemcpy(c->host_inst.data(), c->inst, c->n_inst * sizeof(Inst), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->n_inst; ++i) {
        c->host_inst[i].paused = 0;
        c->host_inst[i].t_pause = t_pause ? t_pause[i] : INFINITY;
    }
    ECHECK(hipMemcpyAsync(c->inst, c->host_inst.data(), c->n_inst * sizeof(Inst), hipMemcpyHostToDevice, c->stream));
    ECHECK(hipStreamSynchronize(c->stream));
    return rebuild_tables(c);
}

extern "C" int hgrp1d_run(hgrp1d *c, int use_graph, double *ms, long long *cell_updates)
{
    if (!c || !c->uploaded) return -1001;
    int rc = arm(c, nullptr);
    if (rc) return rc;
    rc = run_all(c, use_graph, 0, ms);
    if (rc) return rc;
    ECHECK(hipMemcpy(c->host_inst.data(), c->inst, c->n_inst * sizeof(Inst), hipMemcpyDeviceToHost));
    if (cell_updates) {
        long long s = 0;
        for (int i = 0; i < c->n_inst; ++i) {
            const Inst &I = c->host_inst[i];
            // steps that updated the state: k on a normal or stopped end, k - 1 after an abort or the step limit
            const long long upd = (I.status == 1 || I.k > c->P.n_max) ? I.k - 1 : I.k;
            s += upd * c->m[i];
        }
        *cell_updates = s;
    }
    return 0;
}

extern "C" int hgrp1d_run_until(hgrp1d *c, int use_graph, const double *t_pause, int *paused)
{
    if (!c || !c->uploaded || !t_pause) return -1001;
    int rc = arm(c, t_pause);
    if (rc) return rc;
    rc = run_all(c, use_graph, 0, nullptr);
    if (rc) return rc;
    ECHECK(hipMemcpy(c->host_inst.data(), c->inst, c->n_inst * sizeof(Inst), hipMemcpyDeviceToHost));
    if (paused)
        for (int i = 0; i < c->n_inst; ++i) paused[i] = c->host_inst[i].paused && !c->host_inst[i].done;
    return 0;
}

extern "C" int hgrp1d_results(hgrp1d *c, int *steps, int *status, double *time_end)
{
    if (!c || !c->uploaded) return -1001;
    ECHECK(hipMemcpy(c->host_inst.data(), c->inst, c->n_inst * sizeof(Inst), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->n_inst; ++i) {
        const Inst &I = c->host_inst[i];
        if (steps) steps[i] = I.k;
        if (status) status[i] = I.status;
        if (time_end) time_end[i] = I.time_c;
    }
    return 0;
}

extern "C" int hgrp1d_set_blocks(hgrp1d *c, int bs_iface, int bs_update)
{
    if (!c || bs_index(bs_iface) < 0 || bs_index(bs_update) < 0) return -1001;
    c->bs_if = bs_iface;
    c->bs_up = bs_update;
    drop_graph(c);              // captured launches carry the old sizes
    return 0;
}

// Autotuner: times `probe_steps` steps from the uploaded state for every pair
// of block sizes, twice (the faster of the two counts, so one slow sample does
// not decide), keeps the fastest pair, and restores the uploaded state.  A
// multiple of STEPS_PER_GRAPH steps is timed through captured graphs, the way
// hgrp1d_run launches them.  On an error the block sizes and the instances are
// put back as they were.
extern "C" int hgrp1d_tune(hgrp1d *c, int probe_steps, int *bs_iface, int *bs_update)
{
    if (!c || !c->uploaded || probe_steps < 1) return -1001;
    const int use_graph = probe_steps % STEPS_PER_GRAPH == 0 ? 1 : 0;
    const int bi0 = c->bs_if, bu0 = c->bs_up;
    auto fail = [&](int rc) {
        c->bs_if = bi0;
        c->bs_up = bu0;
        (void)reset_instances(c);
        return rc;
    };
    float best = 1e30f;
    int bi = c->bs_if, bu = c->bs_up;
    for (int a : BS_CHOICES)
        for (int b : BS_CHOICES) {
            c->bs_if = a;
            c->bs_up = b;
            double fastest = 1e30;
            for (int trial = 0; trial < 2; ++trial) {
                int rc = reset_instances(c);
                if (rc) return fail(rc);
                double ms = 0.0;
                enqueue_step(c);          // warm the code
                rc = run_all(c, use_graph, probe_steps, &ms);
                if (rc) return fail(rc);
                if (ms < fastest) fastest = ms;
            }
            if (fastest < best) { best = (float)fastest; bi = a; bu = b; }
        }
    c->bs_if = bi;
    c->bs_up = bu;
    if (bs_iface) *bs_iface = bi;
    if (bs_update) *bs_update = bu;
    return reset_instances(c);
}
