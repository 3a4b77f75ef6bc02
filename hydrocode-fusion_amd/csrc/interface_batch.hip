// interface_batch.hip -- batched per-interface entry points (one thread per
// interface) used to check the device GRP solver against the reference's
// per-function vectors.  C ABI in include/hydrocode_grp.h.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/hydrocode_grp.h"
#include "grp_device.hpp"
#include "grp1d_device.hpp"
#include "g2d_device.hpp"

using namespace hgrp;

__global__ void flux_batch_kernel(const double *L, const double *R, int count, double tau, double eps,
                                  int dir, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 15, *r = R + (size_t)k * 15;
    Side SL = {l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]};
    Side SR = {r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
    double F[4], I[4];
    const Gas G = make_gas(l[0]);
    const int code = dir == 0 ? grp_2d_flux<0>(SL, SR, G, tau, eps, F, I, eps)
                              : grp_2d_flux<1>(SL, SR, G, tau, eps, F, I, eps);
    double *o = out + (size_t)k * 9;
    o[0] = code;
    for (int q = 0; q < 4; ++q) { o[1 + q] = F[q]; o[5 + q] = I[q]; }
}

__global__ void riemann_batch_kernel(const double *in, int count, int N, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *a = in + (size_t)k * 10;
    const StarResult r = riemann_exact(make_gas(a[0]), make_gas(a[1]), a[2], a[3], a[4], a[5], a[6], a[7], a[8],
                                       a[9], N);
    double *o = out + (size_t)k * 6;
    o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.gap; o[5] = r.iters;
}

// GRP_2D_flux with the reference's multi-fluid members for the split sweeps'
// normals: sides of 21 doubles (gamma, lambda_u, lambda_v, rho u v p, d_rho
// d_u d_v d_p, t_rho t_u t_v t_p, z_a d_z_a t_z_a, phi d_phi t_phi), of which
// the tangential derivatives and the mass fraction are not read (zero in the
// split path) and gamma follows from z_a (fluids a: ga, b: gb).
// out: code, F_rho F_u F_v F_e, RHO_int U_int V_int P_int, z_a at t + tau.
__global__ void flux_tf_batch_kernel(const double *L, const double *R, int count, double tau, double eps, int dir,
                                     double ga, double gb, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 21, *r = R + (size_t)k * 21;
    Side SL = {l[3], l[4], l[5], l[6], l[7], l[8], l[9], l[10]};
    Side SR = {r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10]};
    TwoGas T;
    T.a = make_gas(ga);
    T.b = make_gas(gb);
    const double FA[4] = {l[15], l[16], r[15], r[16]};
    double F[4], I[4], Z[3];
    const int code = dir == 0 ? grp_2d_flux_tf<0>(SL, SR, T, FA, tau, eps, F, I, Z, eps)
                              : grp_2d_flux_tf<1>(SL, SR, T, FA, tau, eps, F, I, Z, eps);
    double *o = out + (size_t)k * 10;
    o[0] = code;
    for (int q = 0; q < 4; ++q) { o[1 + q] = F[q]; o[5 + q] = I[q]; }
    o[9] = Z[2];
}

// GuessP per row: DL DR UL UR PL PR CL CR GammaL GammaR eps
__global__ void guessp_batch_kernel(const double *in, int count, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *a = in + (size_t)k * 11;
    out[k] = guess_pvrs(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);
}

// The exact solve started from GuessP (at most GUESS_ITERS Newton steps from
// it, then the reference's own guess).  in: rho_L u_L p_L rho_R u_R p_R gamma_L
// gamma_R eps tol; out: u* p* CRW0 CRW1 iterations guess.
__global__ void riemann_pvrs_batch_kernel(const double *in, int count, int N, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *a = in + (size_t)k * 10;
    const double rho_L = a[0], u_L = a[1], p_L = a[2], rho_R = a[3], u_R = a[4], p_R = a[5], gL = a[6], gR = a[7];
    const double c_L = sqrt(gL * p_L / rho_L), c_R = sqrt(gR * p_R / rho_R);
    const double pg = guess_pvrs(rho_L, rho_R, u_L, u_R, p_L, p_R, c_L, c_R, gL, gR, a[8]);
    const StarResult r = riemann_exact(make_gas(gL), make_gas(gR), u_L, u_R, p_L, p_R, c_L, c_R, a[8], a[9], N, pg);
    double *o = out + (size_t)k * 6;
    o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.iters; o[5] = pg;
}

// Interface sides of the per-function entry points: 15 doubles in the order
// gamma, lambda_u, lambda_v, rho, u, v, p, d_rho, d_u, d_v, d_p, t_rho, t_u, t_v, t_p.
__device__ __forceinline__ Side side_of(const double *a)
{
    return Side{a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]};
}

// linear_GRP_solver_Edir_Q1D with every argument of its own (atc, lambda_u/v of
// the left side, a gamma per side).  out: wave_speed[2], D[4], U[4], U_star[6].
__global__ void q1d_batch_kernel(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 15, *r = R + (size_t)k * 15;
    double *o = out + (size_t)k * 16;
    double D[4], U[4], ws[2], us[6];
    const Q1DExtra X = {atc, l[1], l[2], ws, us};
    grp_q1d(D, U, side_of(l), side_of(r), l + 11, r + 11, make_gas(l[0]), make_gas(r[0]), eps, eps, -1.0, nullptr,
            nullptr, &X);
    o[0] = ws[0]; o[1] = ws[1];
    for (int q = 0; q < 4; ++q) { o[2 + q] = D[q]; o[6 + q] = U[q]; }
    for (int q = 0; q < 6; ++q) o[10 + q] = us[q];
}

// linear_GRP_solver_Edir_G2D: sides of 21 doubles (g2_side_of), lambda_u of
// the left side, lambda_v of the right one.  out: wave_speed[2], D[6], U[6],
// U_star[6].
__global__ void g2d_batch_kernel(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 21, *r = R + (size_t)k * 21;
    double *o = out + (size_t)k * 20;
    double D[6], U[6], ws[2], us[6];
    grp_g2d(D, U, ws, us, g2_side_of(l, -1.0), g2_side_of(r, 1.0), l[1], r[2], eps, atc, eps);
    o[0] = ws[0]; o[1] = ws[1];
    for (int q = 0; q < 6; ++q) { o[2 + q] = D[q]; o[8 + q] = U[q]; o[14 + q] = us[q]; }
}

// GRP_2D_flux for any unit normal, both sides' tangential derivatives and
// gammas.  out: code, F_rho F_u F_v F_e, RHO_int U_int V_int P_int (zeros
// when code != 0: the reference returns before it writes them).
__global__ void flux_general_batch_kernel(const double *L, const double *R, int count, double tau, double eps,
                                          double n_x, double n_y, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 15, *r = R + (size_t)k * 15;
    double *o = out + (size_t)k * 9;
    double F[4], I[4];
    const int code = grp_2d_flux_g(side_of(l), side_of(r), l + 11, r + 11, make_gas(l[0]), make_gas(r[0]), n_x, n_y,
                                   tau, eps, F, I, eps);
    o[0] = code;
    for (int q = 0; q < 4; ++q) { o[1 + q] = code ? 0.0 : F[q]; o[5 + q] = code ? 0.0 : I[q]; }
}

// linear_GRP_solver_Edir (1D, one gamma).  out: D[3], U[3] = (rho, u, p).
__global__ void edir_batch_kernel(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *l = L + (size_t)k * 15, *r = R + (size_t)k * 15;
    const Side1 a{l[3], l[4], l[6], l[7], l[8], l[10]}, b{r[3], r[4], r[6], r[7], r[8], r[10]};
    double D[3], U[3];
    grp_edir(D, U, a, b, make_gas(l[0]), eps, atc);
    double *o = out + (size_t)k * 6;
    for (int q = 0; q < 3; ++q) { o[q] = D[q]; o[3 + q] = U[q]; }
}

// plain streaming copy: the calibration kernel for HBM byte counters.  Each
// thread has four 16-byte loads in flight before it stores (grid-stride, every
// index checked against n2), so the copy is limited by HBM, not by latency.
__global__ void __launch_bounds__(256) copy_probe_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                         size_t n2)
{
    const size_t S = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += 4 * S) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u * S < n2) v[u] = src[k + u * S];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u * S < n2) dst[k + u * S] = v[u];
    }
}

__global__ void pow_batch_kernel(const double *x, const double *y, int count, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) out[k] = gpow(x[k], y[k]);
}

namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

int run_batch(const void *const *hin, const size_t *in_bytes, int nin, void *hout, size_t out_bytes,
              void (*launch)(void *const *din, void *dout, void *ctx), void *ctx)
{
    DevBuf bufs[4], ob;
    void *din[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int q = 0; q < nin; ++q) {
        if (hipMalloc(&bufs[q].p, in_bytes[q] ? in_bytes[q] : 8) != hipSuccess) return -2;
        if (in_bytes[q] && hipMemcpy(bufs[q].p, hin[q], in_bytes[q], hipMemcpyHostToDevice) != hipSuccess) return -3;
        din[q] = bufs[q].p;
    }
    if (hipMalloc(&ob.p, out_bytes ? out_bytes : 8) != hipSuccess) return -2;
    launch(din, ob.p, ctx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        fprintf(stderr, "hydrogrp: batch kernel failed: %s\n", hipGetErrorString(e));
        return -(int)e - 1;
    }
    if (out_bytes && hipMemcpy(hout, ob.p, out_bytes, hipMemcpyDeviceToHost) != hipSuccess) return -3;
    return 0;
}

struct FluxCtx { int count; double tau, eps; int dir; };
struct RiemCtx { int count, N; };
struct SideCtx { int count; double a, b, c, d; };   // per-entry-point scalars

int side_batch(const double *L, const double *R, int count, int out_per, double *out, SideCtx sc,
               void (*launch)(void *const *din, void *dout, void *ctx), int width = 15)
{
    if (count < 0 || (count && (!L || !R || !out))) return -1001;
    if (count == 0) return 0;
    const void *in[2] = {L, R};
    const size_t nb[2] = {(size_t)count * width * 8, (size_t)count * width * 8};
    sc.count = count;
    return run_batch(in, nb, 2, out, (size_t)count * out_per * 8, launch, &sc);
}
}  // namespace

extern "C" int hgrp_grp2d_flux_batch(const double *L, const double *R, int count, double tau,
                                     double eps, int dir, double *out)
{
    if (count < 0 || (count && (!L || !R || !out)) || (dir != 0 && dir != 1)) return -1001;
    if (count == 0) return 0;
    const void *in[2] = {L, R};
    const size_t nb[2] = {(size_t)count * 15 * 8, (size_t)count * 15 * 8};
    FluxCtx fc = {count, tau, eps, dir};
    return run_batch(in, nb, 2, out, (size_t)count * 9 * 8,
                     [](void *const *d, void *o, void *cx) {
                         const FluxCtx *f = (const FluxCtx *)cx;
                         hipLaunchKernelGGL(flux_batch_kernel, dim3((f->count + 127) / 128), dim3(128), 0, 0,
                                            (const double *)d[0], (const double *)d[1], f->count, f->tau, f->eps,
                                            f->dir, (double *)o);
                     },
                     &fc);
}

// ---- hgrp_grp2d_flux_batch without allocations ----------------------------------
// The staging of the sides and of the results (what hgrp_grp2d_flux_batch
// allocates per call) lives in one fixed device array, for callers that cannot
// allocate device memory.  Each GPU the library runs on holds its own copy
// until the process ends; every call on a GPU uses that GPU's copy, so calls
// that overlap on one GPU race.  Nothing in it needs clearing: a call writes
// the sides it reads and the results it copies back.
constexpr int NOALLOC_MAX = 1 << 16;                                  // interfaces per call
constexpr size_t NOALLOC_L = 0, NOALLOC_R = (size_t)NOALLOC_MAX * 15, NOALLOC_OUT = (size_t)NOALLOC_MAX * 30;
constexpr size_t NOALLOC_DOUBLES = (size_t)NOALLOC_MAX * (15 + 15 + 9);   // 19.5 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[NOALLOC_DEVICES];   // the address of `scratch` on each device, looked up once

// The current device's copy of `scratch` (its device address: the array's name
// is a host-side stand-in and must not reach a kernel or a copy).
int scratch_here(double **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return -(int)e - 1;
    if (dev < 0 || dev >= NOALLOC_DEVICES) return -1003;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "hydrogrp: no device address for the fixed staging array: %s\n", hipGetErrorString(e));
            return e != hipSuccess ? -(int)e - 1 : -2;
        }
        scratch_dev[dev] = (double *)p;
    }
    *out = scratch_dev[dev];
    return 0;
}
}  // namespace

extern "C" int hgrp_grp2d_flux_batch_noalloc(const double *L, const double *R, int count, double tau, double eps,
                                             int dir, double *out)
{
    if (count < 0 || (count && (!L || !R || !out)) || (dir != 0 && dir != 1)) return -1001;
    if (count > NOALLOC_MAX) {
        fprintf(stderr, "hydrogrp: hgrp_grp2d_flux_batch_noalloc: %d interfaces, at most %d per call\n", count,
                NOALLOC_MAX);
        return -1002;
    }
    if (count == 0) return 0;
    double *base = nullptr;
    const int rc = scratch_here(&base);
    if (rc) return rc;
    double *dL = base + NOALLOC_L, *dR = base + NOALLOC_R, *dO = base + NOALLOC_OUT;
    const size_t nb = (size_t)count * 15 * 8;
    hipError_t e = hipMemcpy(dL, L, nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dR, R, nb, hipMemcpyHostToDevice);
    if (e != hipSuccess) return -(int)e - 1;
    hipLaunchKernelGGL(flux_batch_kernel, dim3((count + 127) / 128), dim3(128), 0, 0, dL, dR, count, tau, eps, dir, dO);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dO, (size_t)count * 9 * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        fprintf(stderr, "hydrogrp: hgrp_grp2d_flux_batch_noalloc: %s\n", hipGetErrorString(e));
        return -(int)e - 1;
    }
    return 0;
}

struct FluxTfCtx { int count; double tau, eps; int dir; double ga, gb; };

extern "C" int hgrp_grp2d_flux_tf_batch(const double *L, const double *R, int count, double tau, double eps, int dir,
                                        double gamma_a, double gamma_b, double *out)
{
    if (count < 0 || (count && (!L || !R || !out)) || (dir != 0 && dir != 1)) return -1001;
    if (count == 0) return 0;
    const void *in[2] = {L, R};
    const size_t nb[2] = {(size_t)count * 21 * 8, (size_t)count * 21 * 8};
    FluxTfCtx fc = {count, tau, eps, dir, gamma_a, gamma_b};
    return run_batch(in, nb, 2, out, (size_t)count * 10 * 8,
                     [](void *const *d, void *o, void *cx) {
                         const FluxTfCtx *f = (const FluxTfCtx *)cx;
                         hipLaunchKernelGGL(flux_tf_batch_kernel, dim3((f->count + 127) / 128), dim3(128), 0, 0,
                                            (const double *)d[0], (const double *)d[1], f->count, f->tau, f->eps,
                                            f->dir, f->ga, f->gb, (double *)o);
                     },
                     &fc);
}

extern "C" int hgrp_riemann_batch(const double *in, int count, int N, double *out)
{
    if (count < 0 || (count && (!in || !out))) return -1001;
    if (count == 0) return 0;
    const void *hin[1] = {in};
    const size_t nb[1] = {(size_t)count * 10 * 8};
    RiemCtx rc = {count, N};
    return run_batch(hin, nb, 1, out, (size_t)count * 6 * 8,
                     [](void *const *d, void *o, void *cx) {
                         const RiemCtx *r = (const RiemCtx *)cx;
                         hipLaunchKernelGGL(riemann_batch_kernel, dim3((r->count + 127) / 128), dim3(128), 0, 0,
                                            (const double *)d[0], r->count, r->N, (double *)o);
                     },
                     &rc);
}

extern "C" int hgrp_q1d_batch(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    return side_batch(L, R, count, 16, out, SideCtx{0, eps, atc, 0.0, 0.0}, [](void *const *d, void *o, void *cx) {
        const SideCtx *c = (const SideCtx *)cx;
        hipLaunchKernelGGL(q1d_batch_kernel, dim3((c->count + 127) / 128), dim3(128), 0, 0, (const double *)d[0],
                           (const double *)d[1], c->count, c->a, c->b, (double *)o);
    });
}

extern "C" int hgrp_g2d_batch(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    return side_batch(
        L, R, count, 20, out, SideCtx{0, eps, atc, 0.0, 0.0},
        [](void *const *d, void *o, void *cx) {
            const SideCtx *c = (const SideCtx *)cx;
            hipLaunchKernelGGL(g2d_batch_kernel, dim3((c->count + 127) / 128), dim3(128), 0, 0, (const double *)d[0],
                               (const double *)d[1], c->count, c->a, c->b, (double *)o);
        },
        21);
}

extern "C" int hgrp_grp2d_flux_general_batch(const double *L, const double *R, int count, double tau, double eps,
                                             double n_x, double n_y, double *out)
{
    return side_batch(L, R, count, 9, out, SideCtx{0, tau, eps, n_x, n_y}, [](void *const *d, void *o, void *cx) {
        const SideCtx *c = (const SideCtx *)cx;
        hipLaunchKernelGGL(flux_general_batch_kernel, dim3((c->count + 127) / 128), dim3(128), 0, 0,
                           (const double *)d[0], (const double *)d[1], c->count, c->a, c->b, c->c, c->d, (double *)o);
    });
}

extern "C" int hgrp_edir_batch(const double *L, const double *R, int count, double eps, double atc, double *out)
{
    return side_batch(L, R, count, 6, out, SideCtx{0, eps, atc, 0.0, 0.0}, [](void *const *d, void *o, void *cx) {
        const SideCtx *c = (const SideCtx *)cx;
        hipLaunchKernelGGL(edir_batch_kernel, dim3((c->count + 127) / 128), dim3(128), 0, 0, (const double *)d[0],
                           (const double *)d[1], c->count, c->a, c->b, (double *)o);
    });
}

extern "C" int hgrp_guessp_batch(const double *in, int count, double *out)
{
    if (count < 0 || (count && (!in || !out))) return -1001;
    if (count == 0) return 0;
    const void *hin[1] = {in};
    const size_t nb[1] = {(size_t)count * 11 * 8};
    int cnt = count;
    return run_batch(hin, nb, 1, out, (size_t)count * 8,
                     [](void *const *d, void *o, void *cx) {
                         const int c = *(const int *)cx;
                         hipLaunchKernelGGL(guessp_batch_kernel, dim3((c + 127) / 128), dim3(128), 0, 0,
                                            (const double *)d[0], c, (double *)o);
                     },
                     &cnt);
}

extern "C" int hgrp_riemann_pvrs_batch(const double *in, int count, int N, double *out)
{
    if (count < 0 || (count && (!in || !out))) return -1001;
    if (count == 0) return 0;
    const void *hin[1] = {in};
    const size_t nb[1] = {(size_t)count * 10 * 8};
    RiemCtx rc = {count, N};
    return run_batch(hin, nb, 1, out, (size_t)count * 6 * 8,
                     [](void *const *d, void *o, void *cx) {
                         const RiemCtx *r = (const RiemCtx *)cx;
                         hipLaunchKernelGGL(riemann_pvrs_batch_kernel, dim3((r->count + 127) / 128), dim3(128), 0, 0,
                                            (const double *)d[0], r->count, r->N, (double *)o);
                     },
                     &rc);
}

extern "C" int hgrp_pow_batch(const double *x, const double *y, int count, double *out)
{
    if (count < 0 || (count && (!x || !y || !out))) return -1001;
    if (count == 0) return 0;
    const void *hin[2] = {x, y};
    const size_t nb[2] = {(size_t)count * 8, (size_t)count * 8};
    int cnt = count;
    return run_batch(hin, nb, 2, out, (size_t)count * 8,
                     [](void *const *d, void *o, void *cx) {
                         const int c = *(const int *)cx;
                         hipLaunchKernelGGL(pow_batch_kernel, dim3((c + 255) / 256), dim3(256), 0, 0,
                                            (const double *)d[0], (const double *)d[1], c, (double *)o);
                     },
                     &cnt);
}

extern "C" int hgrp_copy_probe(size_t bytes, int reps, double *ms_per_copy)
{
    if (bytes < 16 || reps < 1) return -1001;
    const size_t n2 = bytes / 16;
    DevBuf a, b;
    if (hipMalloc(&a.p, n2 * 16) != hipSuccess || hipMalloc(&b.p, n2 * 16) != hipSuccess) return -2;
    if (hipMemset(a.p, 0, n2 * 16) != hipSuccess) return -3;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -4;
    const unsigned blocks = 256 * 8;
    hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, 0, (const double2 *)a.p, (double2 *)b.p, n2);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, 0, (const double2 *)a.p, (double2 *)b.p, n2);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return -(int)e - 1;
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return 0;
}

extern "C" int hgrp_device_name(char *buf, int len)
{
    if (!buf || len < 1) return -1001;
    buf[0] = '\0';
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
    snprintf(buf, (size_t)len, "%s%s%s, %d CUs", prop.name, prop.name[0] ? " " : "", prop.gcnArchName,
             prop.multiProcessorCount);
    return 0;
}
