// sweep2d_tf.hip -- the directional sweeps for two gamma-law fluids.
//
// Same schedule, limiter, reconstruction, GRP solve and update as sweep2d.hip,
// with a fifth cell quantity: the volume fraction a of fluid a.  It is limited
// and reconstructed like the others, carried through the GRP solver with the
// contact (grp_2d_flux_tf), and advanced in quasi-conservative form
//     a_new = a - nu ((u_n a)_r - (u_n a)_l - a ((u_n)_r - (u_n)_l))
// with u_n and a at t + tau/2 on the interfaces.  Because 1/(gamma-1) is
// linear in a and the energy flux uses the same interface values, a uniform
// pressure and velocity across a material interface stay uniform.  In cells
// with a = 1 (or 0) every operation is the single-fluid one, bit for bit.
// Boundary conditions: -2, -4, -7, -24 (the initial-value ghosts of -1 are not
// kept for a).
#include <hip/hip_runtime.h>
#include "grp_device.hpp"
#include "sweep2d.hpp"

namespace hgrp {

constexpr int NQ = 5;   // rho u v p a

template <int DIR>
struct LineTF {
    const SweepArgs &a;
    size_t base, stride;
    int cnt;
    static constexpr int NV = DIR == 0 ? 1 : 2;

    __device__ LineTF(const SweepArgs &args, int ln) : a(args)
    {
        if (DIR == 0) { base = (size_t)ln; stride = (size_t)args.ld; cnt = args.m; }
        else          { base = (size_t)ln * (size_t)args.ld; stride = 1; cnt = args.n; }
    }
    __device__ size_t at(int k) const { return base + (size_t)k * stride; }
    __device__ void cell(int k, double w[NQ]) const
    {
        const size_t o = at(k);
        const double *f = a.st + o;
        w[0] = f[0]; w[1] = f[a.plane]; w[2] = f[2 * a.plane]; w[3] = f[3 * a.plane]; w[4] = a.fa[o];
    }
    __device__ void value(int k, double w[NQ]) const
    {
        if (k >= 0 && k < cnt) { cell(k, w); return; }
        const bool lo = k < 0;
        switch (a.bc) {
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
        default:
            cell(lo ? 0 : cnt - 1, w);
            return;
        }
    }
    __device__ void prev_slopes(int k, double sp[NQ]) const
    {
        const size_t o = at(k);
        for (int q = 0; q < 4; ++q) sp[q] = a.limiter3 ? a.s[o + q * a.plane] : 0.0;
        sp[4] = a.limiter3 ? a.sa[o] : 0.0;
    }
    __device__ void limited_sp(const double vm[NQ], const double v0[NQ], const double vp[NQ], const double sp[NQ],
                               double s[NQ]) const
    {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double s_L = div_by(v0[q] - vm[q], a.div_h);
            const double s_R = div_by(vp[q] - v0[q], a.div_h);
            s[q] = a.limiter3 ? minmod3(a.alpha * s_L, a.alpha * s_R, sp[q]) : minmod2(s_L, s_R);
        }
    }
    __device__ void limited_at(int k, double s[NQ]) const
    {
        double vm[NQ], v0[NQ], vp[NQ], sp[NQ];
        value(k - 1, vm); value(k, v0); value(k + 1, vp);
        prev_slopes(k, sp);
        limited_sp(vm, v0, vp, sp, s);
    }
    // ghost slopes as in sweep2d.hip; the volume fraction is a scalar like rho
    __device__ void ghost_slope(bool lo, double s[NQ]) const
    {
        for (int q = 0; q < NQ; ++q) s[q] = 0.0;
        double t[NQ];
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
    // update of cell k: the single-fluid update with the mixture's gamma, and
    // the volume fraction in quasi-conservative form
    __device__ void update(int k, const double w[NQ], double e, const double Fl[4], const double Fr[4],
                           const double Zl[3], const double Zr[3], const double Il[4], const double Ir[4]) const
    {
        const size_t o = at(k);
        const double nu = a.nu;
        const double mom_x = w[0] * w[1] - nu * (Fr[1] - Fl[1]);
        const double mom_y = w[0] * w[2] - nu * (Fr[2] - Fl[2]);
        const double ene = w[0] * e - nu * (Fr[3] - Fl[3]);
        const double rho = w[0] + -nu * (Fr[0] - Fl[0]);
        const double u = mom_x / rho;
        const double v = mom_y / rho;
        double fa = w[4] - nu * ((Zr[0] - Zl[0]) - w[4] * (Zr[1] - Zl[1]));
        fa = fmin(fmax(fa, 0.0), 1.0);
        const double X = ene - 0.5 * mom_x * u - 0.5 * mom_y * v;
        double p;
        if (fa == 1.0) p = X * a.two.a.gm1();
        else if (fa == 0.0) p = X * a.two.b.gm1();
        else p = X / (fa / (a.two.a.g - 1.0) + (1.0 - fa) / (a.two.b.g - 1.0));
        double *f = a.st_o + o;
        const size_t P = a.plane;
        f[0] = rho; f[P] = u; f[2 * P] = v; f[4 * P] = ene / rho; f[3 * P] = p; a.fa_o[o] = fa;
        if (p < a.eps || rho < a.eps) atomicOr(&a.status->flags, 1u);
#pragma unroll
        for (int q = 0; q < 4; ++q) a.s_o[o + q * P] = div_by(Ir[q] - Il[q], a.div_h);
        a.sa_o[o] = div_by(Zr[2] - Zl[2], a.div_h);
    }
};

__device__ __forceinline__ void recon_tf(const double w[NQ], const double s[NQ], double hh, bool left, Side &f,
                                         double &fa, double &d_fa)
{
    if (left) {
        f.rho = w[0] + hh * s[0]; f.u = w[1] + hh * s[1]; f.v = w[2] + hh * s[2]; f.p = w[3] + hh * s[3];
        fa = w[4] + hh * s[4];
    } else {
        f.rho = w[0] - hh * s[0]; f.u = w[1] - hh * s[1]; f.v = w[2] - hh * s[2]; f.p = w[3] - hh * s[3];
        fa = w[4] - hh * s[4];
    }
    f.d_rho = s[0]; f.d_u = s[1]; f.d_v = s[2]; f.d_p = s[3];
    d_fa = s[4];
}

template <int DIR, int NM>
__device__ __forceinline__ void solve_interface_tf(const SweepArgs &a, const double wl[NQ], const double sl[NQ],
                                                   const double wr[NQ], const double sr[NQ],
                                                   unsigned long long loop_index, double F[4], double I[4],
                                                   double Z[3], double p_guess = -1.0)
{
    Side SL, SR;
    double FA[4];
    recon_tf(wl, sl, a.hh, true, SL, FA[0], FA[1]);
    recon_tf(wr, sr, a.hh, false, SR, FA[2], FA[3]);
    int code = ifvar_bad(SL, SR, a.eps) ? 1 : 0;
    if (grp_2d_flux_tf<DIR, NM != NM_REF>(SL, SR, a.two, FA, a.tau, a.eps, F, I, Z, a.tol, p_guess)) code = 2;
    if (code) atomicMax(&a.status->fail_key, ((loop_index << 2) | (unsigned long long)code) + 1ull);
}

// learned starting pressure for the interface (wl, sl | wr, sr), every lane of
// the wave (uniform call); gamma of each side from its reconstructed fraction
template <int DIR>
__device__ __forceinline__ double interface_guess_tf(const SweepArgs &a, bool active, const double wl[NQ],
                                                     const double sl[NQ], const double wr[NQ], const double sr[NQ],
                                                     uint32_t *lds)
{
    Side SL, SR;
    double aL, daL, aR, daR;
    recon_tf(wl, sl, a.hh, true, SL, aL, daL);
    recon_tf(wr, sr, a.hh, false, SR, aR, daR);
    const double uL = DIR == 0 ? SL.u : SL.v, uR = DIR == 0 ? SR.u : SR.v;
    const double dist = sqrt((SL.rho - SR.rho) * (SL.rho - SR.rho) + (uL - uR) * (uL - uR)
                             + (SL.p - SR.p) * (SL.p - SR.p));
    // Newton runs only past the acoustic shortcut and the equal-state exit of the exact solver
    const bool want = active && !(dist < a.eps) && !(fabs(uL - uR) < a.tol && fabs(SL.p - SR.p) < a.tol);
    const double gL = aL == 1.0 ? a.two.a.g : (aL == 0.0 ? a.two.b.g : mix_gamma(aL, a.two));
    const double gR = aR == 1.0 ? a.two.a.g : (aR == 0.0 ? a.two.b.g : mix_gamma(aR, a.two));
    return learned_guess_wave(a.mlp, want, SL.rho, uL, SL.p, SR.rho, uR, SR.p, gL, gR, lds);
}

__device__ __forceinline__ bool tf_halted(const SweepArgs &a)
{
    if (a.prev == nullptr) return false;
    const SweepStatus pv = *a.prev;
    if (pv.fail_key != 0ull || pv.flags != 0u) {
        if (threadIdx.x == 0) atomicOr(&a.status->flags, 2u);
        return true;
    }
    return false;
}

// X sweep, row-parallel (the layout of sweep_x_rows_kernel with five quantities):
// W waves per block, W - 1 cell rows updated; one block fills a CU at W / 4
// waves per SIMD
template <int W, int NM>
__global__ void __launch_bounds__(64 * W, (W + 3) / 4) sweep_x_tf_kernel(SweepArgs a)
{
    constexpr bool LEARNED = NM == NM_LEARNED;
    __shared__ double xa[W][NQ][64];   // values
    __shared__ double xb[W][NQ][64];   // slopes
    __shared__ double xf[W][4][64];    // fluxes
    __shared__ double xz[W][3][64];    // u_n a, u_n at t + tau/2; a at t + tau
    __shared__ double xi[W][4][64];    // interface values at t + tau
    __shared__ double xe[W][64];       // E of the owned row
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const bool col = i < a.n;
    const int m = a.m;
    const int j0 = blockIdx.y * (W - 1);
    const int c = j0 - 1 + w;
    const LineTF<0> L(a, col ? i : 0);

    const bool own = col && c >= 0 && c < m;
    double v[NQ] = {1.0, 0.0, 0.0, 1.0, 1.0}, vl[NQ], vr[NQ], sp[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0}, e_c = 0.0;
    // every global load before the first LDS store: the owned cell, then the
    // end waves' extra cells
    if (col && c <= m) L.value(c, v);
    if (own) {
        L.prev_slopes(c, sp);
        e_c = a.st[L.at(c) + 4 * a.plane];
    }
    if (col && w == 0 && c >= 0) L.value(c - 1, vl);
    double vrr[NQ], sp1[NQ];
    const bool next_own = col && w == W - 1 && c + 1 < m;
    if (col && w == W - 1 && c + 1 <= m) L.value(c + 1, vr);
    if (next_own) {
        L.value(c + 2, vrr);
        L.prev_slopes(c + 1, sp1);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) xa[w][q][lane] = v[q];
    xe[w][lane] = e_c;
    // the predecessor's status is read while the loads above are in flight
    // (they stay in bounds either way); the whole block returns together
    if (tf_halted(a)) return;
    __syncthreads();
    if (w > 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) vl[q] = xa[w - 1][q][lane];
    if (w < W - 1)
#pragma unroll
        for (int q = 0; q < NQ; ++q) vr[q] = xa[w + 1][q][lane];

    double s[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0}, sr[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (col) {
        if (own) L.limited_sp(vl, v, vr, sp, s);
        else if (c == -1) L.ghost_slope(true, s);
        else if (c == m) L.ghost_slope(false, s);
        if (w == W - 1) {
            if (next_own) L.limited_sp(v, vr, vrr, sp1, sr);
            else if (c + 1 == m) L.ghost_slope(false, sr);
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) xb[w][q][lane] = s[q];
    __syncthreads();
    if (w < W - 1)
#pragma unroll
        for (int q = 0; q < NQ; ++q) sr[q] = xb[w + 1][q][lane];

    double F[4] = {0.0, 0.0, 0.0, 0.0}, I[4] = {0.0, 0.0, 0.0, 0.0}, Z[3] = {0.0, 0.0, 0.0};
    const int f = c + 1;
    double pg = -1.0;
    if constexpr (LEARNED)   // xf[w] is not written before the solve: it holds the network inputs
        pg = interface_guess_tf<0>(a, col && f <= m, v, s, vr, sr, (uint32_t *)&xf[w][0][0]);
    if (col && f <= m)
        solve_interface_tf<0, NM>(a, v, s, vr, sr, (unsigned long long)i * (unsigned long long)(m + 1) + (unsigned long long)f,
                              F, I, Z, pg);
#pragma unroll
    for (int q = 0; q < 4; ++q) { xf[w][q][lane] = F[q]; xi[w][q][lane] = I[q]; }
#pragma unroll
    for (int q = 0; q < 3; ++q) xz[w][q][lane] = Z[q];
    __syncthreads();
    if (own && w >= 1) {
        double Fl[4], Il[4], Zl[3], vc[NQ];
#pragma unroll
        for (int q = 0; q < 4; ++q) { Fl[q] = xf[w - 1][q][lane]; Il[q] = xi[w - 1][q][lane]; }
#pragma unroll
        for (int q = 0; q < 3; ++q) Zl[q] = xz[w - 1][q][lane];
#pragma unroll
        for (int q = 0; q < NQ; ++q) vc[q] = xa[w][q][lane];
        const SweepArgs ar = reload_args<SweepArgs>();   // not held across the solve
        LineTF<0>(ar, i).update(c, vc, xe[w][lane], Fl, F, Zl, Z, Il, I);
    }
}

// Y sweep: a wave owns 64 consecutive cells of a y-line (as sweep_y_kernel)
template <int NM, int G = 16>
__global__ void __launch_bounds__(256, 3) sweep_y_tf_kernel(SweepArgs a, int chunks)
{
    constexpr bool LEARNED = NM == NM_LEARNED;
    // 64/G y-lines of G lanes per wave, as in sweep_y_kernel: lane k of a row
    // owns cell c0-1+k and updates it for k >= 1
    constexpr int ROWS = 64 / G;
    __shared__ double stash[4][NQ + 1][64];
    __shared__ uint32_t mlp_lds[LEARNED ? 4 : 1][LEARNED ? 64 * 8 : 1];
    const int lane = threadIdx.x & 63, k = lane % G;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int jg = (int)(wave / chunks);
    if (jg * ROWS >= a.m) return;
    const int j = jg * ROWS + lane / G;
    const bool rowv = j < a.m;
    const int c0 = (int)(wave % chunks) * (G - 1);
    const int n = a.n;
    const int c = c0 - 1 + k;
    const LineTF<1> L(a, rowv ? j : 0);

    const bool own = rowv && c >= 0 && c < n;
    double w[NQ] = {1.0, 0.0, 0.0, 1.0, 1.0}, sp[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0}, e_c = 0.0;
    if (rowv && c <= n) L.value(c, w);
    if (own) {
        L.prev_slopes(c, sp);
        e_c = a.st[L.at(c) + 4 * a.plane];
    }
    const bool first = rowv && k == 0 && c >= 0, last = rowv && k == G - 1 && c + 1 <= n;
    const bool last_own = rowv && k == G - 1 && c + 1 < n;
    double wl_e[NQ], wr_e[NQ], wrr[NQ], sp1[NQ];
    if (first) L.value(c - 1, wl_e);
    if (last) L.value(c + 1, wr_e);
    if (last_own) {
        L.value(c + 2, wrr);
        L.prev_slopes(c + 1, sp1);
    }
    // the predecessor's status is read while the loads above are in flight
    if (tf_halted(a)) return;
    double(*st)[64] = stash[(threadIdx.x >> 6) & 3];
#pragma unroll
    for (int q = 0; q < NQ; ++q) st[q][lane] = w[q];
    st[NQ][lane] = e_c;
    double wl[NQ], wr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) { wl[q] = __shfl_up(w[q], 1); wr[q] = __shfl_down(w[q], 1); }
    if (first)
#pragma unroll
        for (int q = 0; q < NQ; ++q) wl[q] = wl_e[q];
    if (last)
#pragma unroll
        for (int q = 0; q < NQ; ++q) wr[q] = wr_e[q];

    double s[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (own) L.limited_sp(wl, w, wr, sp, s);
    else if (rowv && c == -1) L.ghost_slope(true, s);
    else if (rowv && c == n) L.ghost_slope(false, s);
    double sr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) sr[q] = __shfl_down(s[q], 1);
    if (rowv && k == G - 1) {
        if (last_own) L.limited_sp(w, wr, wrr, sp1, sr);
        else if (c + 1 == n) L.ghost_slope(false, sr);
    }

    double F[4] = {0.0, 0.0, 0.0, 0.0}, I[4] = {0.0, 0.0, 0.0, 0.0}, Z[3] = {0.0, 0.0, 0.0};
    const int f = c + 1;
    const bool face = rowv && f <= n;
    double pg = -1.0;
    if constexpr (LEARNED) pg = interface_guess_tf<1>(a, face, w, s, wr, sr, mlp_lds[(threadIdx.x >> 6) & 3]);
    if (face)
        solve_interface_tf<1, NM>(a, w, s, wr, sr, (unsigned long long)j * (unsigned long long)(n + 1) + (unsigned long long)f,
                              F, I, Z, pg);
    double Fl[4], Il[4], Zl[3];
#pragma unroll
    for (int q = 0; q < 4; ++q) { Fl[q] = __shfl_up(F[q], 1); Il[q] = __shfl_up(I[q], 1); }
#pragma unroll
    for (int q = 0; q < 3; ++q) Zl[q] = __shfl_up(Z[q], 1);
    if (k >= 1 && own) {
        double wc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) wc[q] = st[q][lane];
        const SweepArgs ar = reload_args<SweepArgs>();   // not held across the solve
        LineTF<1>(ar, j).update(c, wc, st[NQ][lane], Fl, F, Zl, Z, Il, I);
    }
}

// min over cells of min(h_x, h_y) / (|c| + |u| + |v|) with the mixture's gamma
__global__ void __launch_bounds__(256) cfl_tf_kernel(const double *rho, const double *u, const double *v,
                                                     const double *p, const double *fa, int m, int n, int ld,
                                                     TwoGas two, double hmin, unsigned long long *out)
{
    double best = INFINITY;
    const long total = (long)m * n;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const long jj = k / n, ii = k % n;
        const size_t o = (size_t)jj * ld + ii;
        const double g = gas_of(fa[o], two).g;
        const double c = sqrt(g * p[o] / rho[o]);
        const double sigma = fabs(c) + fabs(u[o]) + fabs(v[o]);
        best = fmin(best, hmin / sigma);
    }
    for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_xor(best, off));
    if ((threadIdx.x & 63) == 0) {
        unsigned long long bits = __double_as_longlong(best);
        if (!(best >= 0.0)) bits = 0x7ff0000000000000ull;
        atomicMin(out, bits);
    }
}

// E of every cell from its primitives with the mixture's gamma: the reference
// main program's expression (src/hydrocode_2D/hydrocode.c:251-252) in pure
// cells, and p * (a/(gamma_a-1) + (1-a)/(gamma_b-1)) / rho in mixed ones, the
// exact inverse of the pressure the update forms there.
__global__ void __launch_bounds__(256) energy_tf_kernel(const double *rho, const double *u, const double *v,
                                                        const double *p, const double *fa, double *e, int m, int n,
                                                        int ld, TwoGas two)
{
    const long total = (long)m * n;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const long jj = k / n, ii = k % n;
        const size_t o = (size_t)jj * ld + ii;
        const double a = fa[o];
        double E;
        if (a == 1.0 || a == 0.0) {
            const double g = a == 1.0 ? two.a.g : two.b.g;
            E = 0.5 * u[o] * u[o] + p[o] / (g - 1.0) / rho[o];
        } else {
            E = 0.5 * u[o] * u[o] + p[o] * (a / (two.a.g - 1.0) + (1.0 - a) / (two.b.g - 1.0)) / rho[o];
        }
        e[o] = E + 0.5 * v[o] * v[o];
    }
}

hipError_t launch_energy_tf(const double *rho, const double *u, const double *v, const double *p, const double *fa,
                            double *e, int m, int n, int ld, const TwoGas &two, hipStream_t st)
{
    const long total = (long)m * n;
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(energy_tf_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rho, u, v, p, fa, e, m, n, ld, two);
    return hipGetLastError();
}

// the two-fluid sweep kernel's name as profilers print it
const char *sweep_tf_kernel_name(int dir, int nm)
{
    static const char *xn[3] = {"sweep_x_tf_kernel<12, 0>", "sweep_x_tf_kernel<12, 1>", "sweep_x_tf_kernel<12, 2>"};
    static const char *yn[3] = {"sweep_y_tf_kernel<0, 16>", "sweep_y_tf_kernel<1, 16>", "sweep_y_tf_kernel<2, 16>"};
    if (nm < 0 || nm > 2 || dir < 0 || dir > 1) return "";
    return dir == 0 ? xn[nm] : yn[nm];
}

// Rows per block of the X sweep: 12 waves (11 rows updated) at 3 waves per SIMD,
// 132 KiB of LDS
constexpr int TF_X_W = 12;

hipError_t launch_sweep_x_tf(const SweepArgs &a, hipStream_t st)
{
    constexpr int W = TF_X_W;
    dim3 grid((a.n + 63) / 64, (a.m + W - 2) / (W - 1));
    switch (newton_mode(a)) {
    case NM_LEARNED: hipLaunchKernelGGL((sweep_x_tf_kernel<W, NM_LEARNED>), grid, dim3(64 * W), 0, st, a); break;
    case NM_CONV: hipLaunchKernelGGL((sweep_x_tf_kernel<W, NM_CONV>), grid, dim3(64 * W), 0, st, a); break;
    default: hipLaunchKernelGGL((sweep_x_tf_kernel<W, NM_REF>), grid, dim3(64 * W), 0, st, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_sweep_y_tf(const SweepArgs &a, hipStream_t st)
{
    // 4 rows of 16 lanes per wave (the default of the single-fluid Y sweep)
    const int chunks = (a.n + 14) / 15;
    const long waves = (long)((a.m + 3) / 4) * chunks;
    const long blocks = (waves + 3) / 4;
    switch (newton_mode(a)) {
    case NM_LEARNED: hipLaunchKernelGGL(sweep_y_tf_kernel<NM_LEARNED>, dim3((unsigned)blocks), dim3(256), 0, st, a, chunks); break;
    case NM_CONV: hipLaunchKernelGGL(sweep_y_tf_kernel<NM_CONV>, dim3((unsigned)blocks), dim3(256), 0, st, a, chunks); break;
    default: hipLaunchKernelGGL(sweep_y_tf_kernel<NM_REF>, dim3((unsigned)blocks), dim3(256), 0, st, a, chunks); break;
    }
    return hipGetLastError();
}

hipError_t launch_cfl_tf(const double *rho, const double *u, const double *v, const double *p, const double *fa,
                         int m, int n, int ld, const TwoGas &two, double hmin, unsigned long long *out, hipStream_t st)
{
    const long total = (long)m * n;
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(cfl_tf_kernel, dim3((unsigned)blocks), dim3(256), 0, st, rho, u, v, p, fa, m, n, ld, two, hmin,
                       out);
    return hipGetLastError();
}

}  // namespace hgrp
