// sweep2d.hip -- gfx950 kernels of the dimension-split 2D GRP step.
//
//   sweep_x_rows_kernel  one X(tau_s) sweep: a block of W waves covers 64
//                    consecutive i (one coalesced 512-byte row segment per wave)
//                    and W cell rows; x-neighbours are exchanged through LDS.
//   sweep_y_kernel   one Y(tau) sweep: a wave holds 64/G rows of G consecutive
//                    cells of a y-line (contiguous in memory) and exchanges
//                    neighbour values, slopes and fluxes with lane shuffles.
// Each kernel takes the Newton mode (NM_*) as a template argument.
//   cfl_kernel       h/S_max over the frozen initial data (order-independent min).
//
// Semantics follow src/finite_volume/grp_solver_2D_split_EUL_source.c:182-303,
// src/inter_process/bound_cond_slope_limiter_{x,y}.c:27-129,
// src/inter_process/slope_limiter{,_2D_x}.c and src/flux_calc/flux_generator_{x,y}.c:31-165.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include "grp_device.hpp"
#include "sweep2d.hpp"
#include "sweep2d_body.hpp"

namespace hgrp {

// X sweep, row-parallel form (line_sweep<0>): lanes are 64 consecutive i,
// waves the cell rows j0-1 .. j0+W-2.
template <int W, int MINW, int NM>
__global__ void __launch_bounds__(64 * W, MINW) sweep_x_rows_kernel(SweepArgs a)
{
    __shared__ double xa[W][4][64];   // values
    __shared__ double xb[W][4][64];   // slopes
    __shared__ double xf[W][4][64];   // fluxes
    __shared__ double xi[W][4][64];   // interface values at t + tau
    __shared__ double xe[W][64];      // E of the owned cell (read back for the update)
    line_sweep<0, W, NM>(a, KernargArgs{}, blockIdx.x, blockIdx.y, xa, xb, xf, xi, xe);
}

// X sweep walking K chunks of W cell rows per block (line_sweep with K > 1).
template <int W, int MINW, int NM, int K>
__global__ void __launch_bounds__(64 * W, MINW) sweep_x_chunks_kernel(SweepArgs a)
{
    __shared__ double xa[W][4][64];
    __shared__ double xb[W][4][64];
    __shared__ double xf[W][4][64];
    __shared__ double xi[W][4][64];
    __shared__ double xe[W][64];
    __shared__ double xc[2][8][64];   // the last face of the previous chunk: F, I
    (void)a;                          // read through KernargArgs, chunk by chunk
    line_sweep_chunks<0, W, NM, K>(KernargArgs{}, blockIdx.x, blockIdx.y, xa, xb, xf, xi, xe, xc);
}

// Y sweep (y_sweep): blocks of 4 waves, wave index = global wave number.
template <int MINW, int NM, int G = 64>
__global__ void __launch_bounds__(256, MINW) sweep_y_kernel(SweepArgs a, int chunks)
{
    constexpr bool LEARNED = NM == NM_LEARNED;
    __shared__ uint32_t mlp_lds[LEARNED ? 4 : 1][LEARNED ? 64 * 8 : 1];
    __shared__ double stash[4][5][64];   // each wave's cell values and E, read back for the update
    const int wv = (threadIdx.x >> 6) & 3;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    y_sweep<NM, G>(a, KernargArgs{}, wave, chunks, stash[wv], mlp_lds[wv]);
}

// min over cells of min(h_x, h_y) / (|c| + |u| + |v|)  (grp_solver_2D_split_EUL_source.c:185-191).
// fmin is exact, so the order of the reduction does not matter; positive
// doubles order like their bit patterns, so a 64-bit atomicMin reduces them.
__global__ void __launch_bounds__(256) cfl_kernel(const double *rho, const double *u, const double *v,
                                                  const double *p, int m, int n, int ld, double gamma,
                                                  double hmin, unsigned long long *out)
{
    double best = INFINITY;
    const long total = (long)m * n;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const long j = k / n, i = k % n;
        const size_t o = (size_t)j * ld + i;
        const double c = sqrt(gamma * p[o] / rho[o]);
        const double sigma = fabs(c) + fabs(u[o]) + fabs(v[o]);
        best = fmin(best, hmin / sigma);
    }
    for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_xor(best, off));
    if ((threadIdx.x & 63) == 0) {
        unsigned long long bits = __double_as_longlong(best);
        if (!(best >= 0.0)) bits = 0x7ff0000000000000ull;   // NaN never wins a fmin
        atomicMin(out, bits);
    }
}

// ---- launchers -----------------------------------------------------------------
// Launch variants of each sweep (the tuner, hgrp2d_tune, times them on the
// context's own data; the defaults, variant 0, were the fastest on four-quadrant
// runs at 1024^2 and 4096^2 on MI355X):
//   X  0: 8 rows per block, register budget for 4 waves/SIMD (128 VGPRs, a few spills)
//      1: 8 rows per block, the compiler's budget (~150 VGPRs, 3 waves/SIMD)
//      2: 16 rows per block, 4 waves/SIMD
//      3, 4: 8 rows per block walking 4, 8 chunks (31, 63 rows per block; 4 waves/SIMD)
//      5: 8 rows per block walking 6 chunks (47 rows per block)
//   Y  0: 4 rows of 16 lanes per wave, 4 waves/SIMD (a front lying along y meets
//         fewer waves, with more lanes on it: 12 % faster than variant 2 at 1024^2)
//      1: 2 rows of 32 lanes;  2: one 64-lane row;  3: 8 rows of 8 lanes (all 4 waves/SIMD)
//      4: one 64-lane row, the compiler's budget
int sweep_variants(int dir) { return dir == 0 ? 6 : 5; }

template <int NM>
static hipError_t launch_x(const SweepArgs &a, int v, hipStream_t st)
{
    switch (v) {
    case 1:
        hipLaunchKernelGGL((sweep_x_rows_kernel<8, 1, NM>), dim3((a.n + 63) / 64, (a.m + 6) / 7), dim3(64 * 8), 0,
                           st, a);
        break;
    case 3:
        hipLaunchKernelGGL((sweep_x_chunks_kernel<8, 4, NM, 4>), dim3((a.n + 63) / 64, (a.m + 30) / 31), dim3(64 * 8),
                           0, st, a);
        break;
    case 4:
        hipLaunchKernelGGL((sweep_x_chunks_kernel<8, 4, NM, 8>), dim3((a.n + 63) / 64, (a.m + 62) / 63), dim3(64 * 8),
                           0, st, a);
        break;
    case 5:
        hipLaunchKernelGGL((sweep_x_chunks_kernel<8, 4, NM, 6>), dim3((a.n + 63) / 64, (a.m + 46) / 47), dim3(64 * 8),
                           0, st, a);
        break;
    case 2:
        hipLaunchKernelGGL((sweep_x_rows_kernel<16, 4, NM>), dim3((a.n + 63) / 64, (a.m + 14) / 15), dim3(64 * 16),
                           0, st, a);
        break;
    default:
        hipLaunchKernelGGL((sweep_x_rows_kernel<8, 4, NM>), dim3((a.n + 63) / 64, (a.m + 6) / 7), dim3(64 * 8), 0,
                           st, a);
        break;
    }
    return hipGetLastError();
}

static int y_lanes(int v) { return v == 1 ? 32 : (v == 2 || v == 4) ? 64 : v == 3 ? 8 : 16; }

template <int NM>
static hipError_t launch_y(const SweepArgs &a, int v, hipStream_t st)
{
    const int G = y_lanes(v), rows = 64 / G;
    const int chunks = (a.n + G - 2) / (G - 1);
    const long waves = (long)((a.m + rows - 1) / rows) * chunks;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    switch (v) {
    case 1: hipLaunchKernelGGL((sweep_y_kernel<4, NM, 32>), grid, block, 0, st, a, chunks); break;
    case 2: hipLaunchKernelGGL((sweep_y_kernel<4, NM, 64>), grid, block, 0, st, a, chunks); break;
    case 3: hipLaunchKernelGGL((sweep_y_kernel<4, NM, 8>), grid, block, 0, st, a, chunks); break;
    case 4: hipLaunchKernelGGL((sweep_y_kernel<1, NM, 64>), grid, block, 0, st, a, chunks); break;
    default: hipLaunchKernelGGL((sweep_y_kernel<4, NM, 16>), grid, block, 0, st, a, chunks); break;
    }
    return hipGetLastError();
}

hipError_t launch_sweep_x(const SweepArgs &a, int variant, hipStream_t st)
{
    switch (newton_mode(a)) {
    case NM_LEARNED: return launch_x<NM_LEARNED>(a, variant, st);
    case NM_CONV: return launch_x<NM_CONV>(a, variant, st);
    default: return launch_x<NM_REF>(a, variant, st);
    }
}

hipError_t launch_sweep_y(const SweepArgs &a, int variant, hipStream_t st)
{
    switch (newton_mode(a)) {
    case NM_LEARNED: return launch_y<NM_LEARNED>(a, variant, st);
    case NM_CONV: return launch_y<NM_CONV>(a, variant, st);
    default: return launch_y<NM_REF>(a, variant, st);
    }
}

// the kernel's name as profilers print it
const char *sweep_kernel_name(int dir, int variant, int nm)
{
    static const char *xn[3][6] = {
        {"sweep_x_rows_kernel<8, 4, 0>", "sweep_x_rows_kernel<8, 1, 0>", "sweep_x_rows_kernel<16, 4, 0>",
         "sweep_x_chunks_kernel<8, 4, 0, 4>", "sweep_x_chunks_kernel<8, 4, 0, 8>", "sweep_x_chunks_kernel<8, 4, 0, 6>"},
        {"sweep_x_rows_kernel<8, 4, 1>", "sweep_x_rows_kernel<8, 1, 1>", "sweep_x_rows_kernel<16, 4, 1>",
         "sweep_x_chunks_kernel<8, 4, 1, 4>", "sweep_x_chunks_kernel<8, 4, 1, 8>", "sweep_x_chunks_kernel<8, 4, 1, 6>"},
        {"sweep_x_rows_kernel<8, 4, 2>", "sweep_x_rows_kernel<8, 1, 2>", "sweep_x_rows_kernel<16, 4, 2>",
         "sweep_x_chunks_kernel<8, 4, 2, 4>", "sweep_x_chunks_kernel<8, 4, 2, 8>", "sweep_x_chunks_kernel<8, 4, 2, 6>"}};
    static const char *yn[3][5] = {
        {"sweep_y_kernel<4, 0, 16>", "sweep_y_kernel<4, 0, 32>", "sweep_y_kernel<4, 0, 64>", "sweep_y_kernel<4, 0, 8>",
         "sweep_y_kernel<1, 0, 64>"},
        {"sweep_y_kernel<4, 1, 16>", "sweep_y_kernel<4, 1, 32>", "sweep_y_kernel<4, 1, 64>", "sweep_y_kernel<4, 1, 8>",
         "sweep_y_kernel<1, 1, 64>"},
        {"sweep_y_kernel<4, 2, 16>", "sweep_y_kernel<4, 2, 32>", "sweep_y_kernel<4, 2, 64>", "sweep_y_kernel<4, 2, 8>",
         "sweep_y_kernel<1, 2, 64>"}};
    if (nm < 0 || nm > 2 || variant < 0 || variant >= sweep_variants(dir)) return "";
    return dir == 0 ? xn[nm][variant] : yn[nm][variant];
}

hipError_t launch_cfl(const double *rho, const double *u, const double *v, const double *p, int m, int n,
                      int ld, double gamma, double hmin, unsigned long long *out, hipStream_t st)
{
    const long total = (long)m * n;
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(cfl_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rho, u, v, p, m, n, ld, gamma, hmin, out);
    return hipGetLastError();
}

}  // namespace hgrp
