// units2d.hip -- the 2D step's stages as functions of their own: the minmod
// slopes of one direction and the interface fluxes of one direction over a
// whole grid, for callers of the reference's bound_cond_slope_limiter_{x,y}
// (src/inter_process/bound_cond_slope_limiter_x.c:27-129, _y.c:27-129) and
// flux_generator_{x,y} (src/flux_calc/flux_generator_x.c:31-165, _y.c:31-165).
// The time step itself never calls these: its kernels fuse the same stages
// without storing slopes or fluxes.  One thread per cell (slopes) or per
// interface (fluxes); host arrays in and out.  C ABI in include/hydrocode_grp.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "../../include/hydrocode_grp.h"
#include "grp_device.hpp"

using namespace hgrp;

namespace {

// value q (rho u v p) of cell (j, i) of an m x n grid, fields [q][j*n + i]
__device__ __forceinline__ double at(const double *f, int q, int j, int i, int m, int n)
{
    return f[(size_t)q * m * n + (size_t)j * n + i];
}

// minmod_limiter_2D_x (src/inter_process/slope_limiter_2D_x.c:32-84) for dir 0,
// minmod_limiter (slope_limiter.c:31-83) along each row for dir 1
__global__ void slopes_kernel(int dir, int m, int n, const double *state, const double *lo, const double *hi,
                              const double *s_prev, int found, double alpha, double h, double *s_out)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)m * n) return;
    const int j = (int)(k / n), i = (int)(k % n);
    const int pos = dir == 0 ? j : i, len = dir == 0 ? m : n, line = dir == 0 ? i : j;
    for (int q = 0; q < 4; ++q) {
        const double u = at(state, q, j, i, m, n);
        const double um = pos ? (dir == 0 ? at(state, q, j - 1, i, m, n) : at(state, q, j, i - 1, m, n))
                              : lo[line * 4 + q];
        const double up = pos < len - 1 ? (dir == 0 ? at(state, q, j + 1, i, m, n) : at(state, q, j, i + 1, m, n))
                                        : hi[line * 4 + q];
        const double s_L = (u - um) / h;
        const double s_R = (up - u) / h;
        const size_t o = (size_t)q * m * n + k;
        s_out[o] = found ? minmod3(alpha * s_L, alpha * s_R, s_prev[o]) : minmod2(s_L, s_R);
    }
}

// One interface of flux_generator_x (dir 0: interface (j, i), j = 0..m, between
// cells j-1 and j of row i) or flux_generator_y (dir 1: interface (j, i),
// i = 0..n, between cells i-1 and i of line j).  sn: the slopes normal to the
// interfaces, st: the other direction's (transversal; negated for y as the
// reference does).  bfv_lo/hi: per line rho u v p, S*, T* (12 doubles).
__global__ void fluxes_kernel(int dir, int m, int n, const double *state, const double *sn, const double *st,
                              const double *bfv_lo, const double *bfv_hi, double tau, double h, double gamma,
                              double eps, int transversal, double *F, double *I, int *codes)
{
    const int nif_line = dir == 0 ? m + 1 : n + 1, lines = dir == 0 ? n : m;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)nif_line * lines) return;
    int j, i;
    size_t out;
    if (dir == 0) {            // reference loop: i outer, j inner; stored at [j][i]
        i = (int)(k / nif_line);
        j = (int)(k % nif_line);
        out = (size_t)j * n + i;
    } else {                   // j outer, i inner; stored at [j][i] of m x (n+1)
        j = (int)(k / nif_line);
        i = (int)(k % nif_line);
        out = (size_t)j * (n + 1) + i;
    }
    const int pos = dir == 0 ? j : i, len = dir == 0 ? m : n, line = dir == 0 ? i : j;
    const double tsign = dir == 0 ? 1.0 : -1.0;
    double vL[4], dL[4], tL[4], vR[4], dR[4], tR[4];
    for (int q = 0; q < 4; ++q) {
        if (pos) {
            const int cj = dir == 0 ? j - 1 : j, ci = dir == 0 ? i : i - 1;
            dL[q] = at(sn, q, cj, ci, m, n);
            vL[q] = at(state, q, cj, ci, m, n) + 0.5 * h * dL[q];
            tL[q] = transversal ? tsign * at(st, q, cj, ci, m, n) : tsign * 0.0;
        } else {
            const double *b = bfv_lo + (size_t)line * 12;
            dL[q] = b[4 + q];
            vL[q] = b[q] + 0.5 * h * b[4 + q];
            tL[q] = transversal ? tsign * b[8 + q] : tsign * 0.0;
        }
        if (pos < len) {
            dR[q] = at(sn, q, j, i, m, n);
            vR[q] = at(state, q, j, i, m, n) - 0.5 * h * dR[q];
            tR[q] = transversal ? tsign * at(st, q, j, i, m, n) : tsign * 0.0;
        } else {
            const double *b = bfv_hi + (size_t)line * 12;
            dR[q] = b[4 + q];
            vR[q] = b[q] - 0.5 * h * b[4 + q];
            tR[q] = transversal ? tsign * b[8 + q] : tsign * 0.0;
        }
    }
    const Side L = {vL[0], vL[1], vL[2], vL[3], dL[0], dL[1], dL[2], dL[3]};
    const Side R = {vR[0], vR[1], vR[2], vR[3], dR[0], dR[1], dR[2], dR[3]};
    const bool bad_in = ifvar_bad_t(L, R, tL, tR, eps);
    const Gas G = make_gas(gamma);
    double Fo[4], Io[4];
    const int code = grp_2d_flux_g(L, R, tL, tR, G, G, dir == 0 ? 1.0 : 0.0, dir == 0 ? 0.0 : 1.0, tau, eps, Fo, Io,
                                   eps);
    const size_t nif = (size_t)nif_line * lines;
    for (int q = 0; q < 4; ++q) {
        F[q * nif + out] = code ? NAN : Fo[q];
        I[q * nif + out] = code ? NAN : Io[q];
    }
    codes[out] = code ? 2 : (bad_in ? 1 : 0);
}

struct DevArr {
    void *p = nullptr;
    ~DevArr() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8); }
};

hipError_t put(DevArr &d, const void *h, size_t bytes)
{
    hipError_t e = d.alloc(bytes);
    if (e == hipSuccess && h && bytes) e = hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice);
    return e;
}

int fail(hipError_t e, const char *what)
{
    fprintf(stderr, "hydrogrp: %s failed: %s\n", what, hipGetErrorString(e));
    return -(int)e - 1;
}

}  // namespace

extern "C" int hgrp2d_slopes(int dir, int m, int n, const double *state, const double *ghost_lo,
                             const double *ghost_hi, const double *s_prev, int found, double alpha, double h,
                             double *s_out)
{
    if ((dir != 0 && dir != 1) || m < 1 || n < 1 || !state || !ghost_lo || !ghost_hi || !s_out ||
        (found && !s_prev) || (size_t)m * n > ((size_t)1 << 31))
        return -1001;
    const size_t mn = (size_t)m * n, nl = dir == 0 ? (size_t)n : (size_t)m;
    DevArr ds, dl, dh, dp, dout;
    hipError_t e = put(ds, state, 4 * mn * 8);
    if (e == hipSuccess) e = put(dl, ghost_lo, nl * 4 * 8);
    if (e == hipSuccess) e = put(dh, ghost_hi, nl * 4 * 8);
    if (e == hipSuccess) e = put(dp, found ? s_prev : nullptr, found ? 4 * mn * 8 : 0);
    if (e == hipSuccess) e = dout.alloc(4 * mn * 8);
    if (e != hipSuccess) return fail(e, "slopes set-up");
    hipLaunchKernelGGL(slopes_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, 0, dir, m, n,
                       (const double *)ds.p, (const double *)dl.p, (const double *)dh.p, (const double *)dp.p,
                       found, alpha, h, (double *)dout.p);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(s_out, dout.p, 4 * mn * 8, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : fail(e, "slopes kernel");
}

extern "C" int hgrp2d_fluxes(int dir, int m, int n, const double *state, const double *slope_n,
                             const double *slope_t, const double *bfv_lo, const double *bfv_hi, double tau, double h,
                             double gamma, double eps, int transversal, double *F, double *I, int *codes)
{
    if ((dir != 0 && dir != 1) || m < 1 || n < 1 || !state || !slope_n || (transversal && !slope_t) || !bfv_lo ||
        !bfv_hi || !F || !I || !codes || (size_t)(m + 1) * (n + 1) > ((size_t)1 << 31))
        return -1001;
    const size_t mn = (size_t)m * n, nl = dir == 0 ? (size_t)n : (size_t)m;
    const size_t nif = dir == 0 ? (size_t)(m + 1) * n : (size_t)m * (n + 1);
    DevArr ds, dn, dt, dl, dh, dF, dI, dc;
    hipError_t e = put(ds, state, 4 * mn * 8);
    if (e == hipSuccess) e = put(dn, slope_n, 4 * mn * 8);
    if (e == hipSuccess) e = put(dt, transversal ? slope_t : nullptr, transversal ? 4 * mn * 8 : 0);
    if (e == hipSuccess) e = put(dl, bfv_lo, nl * 12 * 8);
    if (e == hipSuccess) e = put(dh, bfv_hi, nl * 12 * 8);
    if (e == hipSuccess) e = dF.alloc(4 * nif * 8);
    if (e == hipSuccess) e = dI.alloc(4 * nif * 8);
    if (e == hipSuccess) e = dc.alloc(nif * 4);
    if (e != hipSuccess) return fail(e, "fluxes set-up");
    hipLaunchKernelGGL(fluxes_kernel, dim3((unsigned)((nif + 127) / 128)), dim3(128), 0, 0, dir, m, n,
                       (const double *)ds.p, (const double *)dn.p, (const double *)dt.p, (const double *)dl.p,
                       (const double *)dh.p, tau, h, gamma, eps, transversal, (double *)dF.p, (double *)dI.p,
                       (int *)dc.p);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(F, dF.p, 4 * nif * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(I, dI.p, 4 * nif * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(codes, dc.p, nif * 4, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : fail(e, "fluxes kernel");
}
