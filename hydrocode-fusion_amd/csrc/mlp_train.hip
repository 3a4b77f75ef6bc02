// mlp_train.hip -- the learned p* guess: training data from the exact solver,
// forward/backward and Adam kernels, packing for the MFMA inference, and the
// batched entry points.  C ABI in include/hydrocode_grp.h.
//
// Training numerics are the inference numerics: the forward pass sees BF16
// weights (W1, W2, w3), inputs as the BF16 hi + lo pair, hidden layer 1 rounded
// to BF16, FP32 biases and FP32 accumulation; gradients pass the roundings
// straight through to the FP32 master weights.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/hydrocode_grp.h"
#include "grp_device.hpp"
#include "mlp_guess.hpp"

using namespace hgrp;

#define MCHECK(x)                                                     \
    do {                                                              \
        hipError_t e_ = (x);                                          \
        if (e_ != hipSuccess) {                                       \
            fprintf(stderr, "hydrogrp: %s failed: %s (%s:%d)\n", #x,  \
                    hipGetErrorString(e_), __FILE__, __LINE__);       \
            return -(int)e_ - 1;                                      \
        }                                                             \
    } while (0)

namespace hgrp {

void mlp_pack_host(const float *P, MlpPacked *W)
{
    memset(W, 0, sizeof *W);
    for (int ob = 0; ob < 2; ++ob)
        for (int lane = 0; lane < 64; ++lane) {
            const int row = 16 * ob + (lane & 15), k0 = 4 * (lane >> 4);
            uint32_t b[4];
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + j;
                const float v = k < 6 ? P[MLP_OFF_W1 + row * 6 + k] : (k < 12 ? P[MLP_OFF_W1 + row * 6 + k - 6] : 0.f);
                b[j] = bf16_bits(v);
            }
            W->w1f[ob][lane] = make_uint2(b[0] | (b[1] << 16), b[2] | (b[3] << 16));
            for (int kb = 0; kb < 2; ++kb) {
                for (int j = 0; j < 4; ++j) b[j] = bf16_bits(P[MLP_OFF_W2 + row * 32 + 16 * kb + k0 + j]);
                W->w2f[ob][kb][lane] = make_uint2(b[0] | (b[1] << 16), b[2] | (b[3] << 16));
            }
        }
    for (int o = 0; o < MLP_H; ++o) {
        W->b1[o] = P[MLP_OFF_B1 + o];
        W->b2[o] = P[MLP_OFF_B2 + o];
        W->w3[o] = bf16_value(P[MLP_OFF_W3 + o]);
    }
    W->b3 = P[MLP_OFF_B3];
}

}  // namespace hgrp

// ---- training data ---------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(uint64_t seed, uint64_t idx, int stream)
{
    return (double)(mix64(seed ^ mix64(idx * 4u + (uint64_t)stream)) >> 11) * 0x1.0p-53;
}

// One sample: left state (1, 0, 1), rho_L/rho_R log-uniform in [1e-3, 1e3],
// p_L/p_R log-uniform in [1e-4, 1e4], (u_R - u_L)/c_L uniform in [-5, 5];
// half the samples have gamma 1.4 on both sides, the other half gamma_L and
// gamma_R independently uniform in [1.2, 1.8] (two-fluid interfaces); p* from
// the exact solver (the reference's guess, tolerance tol).  Samples whose
// solve does not converge (above gamma 1.4 a jump of 5 c_L can open a vacuum)
// are marked not ok and left out of training.
// states: rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R p*.
__global__ void mlp_data_kernel(uint64_t seed, int count, double tol, float *feat, float *target, double *states,
                                int *ok)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const bool mixed = unif(seed, (uint64_t)k, 3) < 0.5;
    const double gL = mixed ? 1.2 + 0.6 * unif(seed ^ 0x5bd1e995ull, (uint64_t)k, 0) : 1.4;
    const double gR = mixed ? 1.2 + 0.6 * unif(seed ^ 0x5bd1e995ull, (uint64_t)k, 1) : 1.4;
    const double rho_L = 1.0, p_L = 1.0, u_L = 0.0;
    const double rho_R = rho_L / pow(10.0, -3.0 + 6.0 * unif(seed, (uint64_t)k, 0));
    const double p_R = p_L / pow(10.0, -4.0 + 8.0 * unif(seed, (uint64_t)k, 1));
    const double c_L = sqrt(gL * p_L / rho_L), c_R = sqrt(gR * p_R / rho_R);
    const double u_R = u_L + (-5.0 + 10.0 * unif(seed, (uint64_t)k, 2)) * c_L;
    const StarResult r = riemann_exact<true>(make_gas(gL), make_gas(gR), u_L, u_R, p_L, p_R, c_L, c_R, 1e-9, tol, 500);
    float x[MLP_IN], p_tr;
    const bool good = mlp_features(rho_L, u_L, p_L, rho_R, u_R, p_R, gL, gR, x, &p_tr) && r.p_star > 0.0 &&
                      isfinite(r.p_star) && r.iters < 500;
    for (int q = 0; q < MLP_IN; ++q) feat[(size_t)k * MLP_IN + q] = good ? x[q] : 0.f;
    target[k] = good ? (float)log(r.p_star / (double)p_tr) : 0.f;
    ok[k] = good;
    if (states) {
        double *s = states + (size_t)k * 9;
        s[0] = rho_L; s[1] = u_L; s[2] = p_L; s[3] = rho_R; s[4] = u_R; s[5] = p_R; s[6] = gL; s[7] = gR;
        s[8] = r.p_star;
    }
}

// ---- forward / backward ------------------------------------------------------------
constexpr int TB = 64;                    // samples per block (one wave)
constexpr int PSTRIDE = MLP_NP + 1;       // partial gradients + loss per block

__device__ __forceinline__ float gelu_grad(float x)
{
    const float k = 0.7978845608f, c = 0.044715f;
    const float a = k * (x + c * x * x * x);
    const float s = 1.0f / (1.0f + __expf(-2.0f * a));
    return s + x * s * (1.0f - s) * 2.0f * k * (1.0f + 3.0f * c * x * x);
}

// Block b handles samples start + b*TB .. +TB-1 (those < count with ok set);
// writes the gradient of sum_s (y_s - t_s)^2 * inv_b and the sum of squared
// errors to partial[b].
__global__ void __launch_bounds__(TB) mlp_fwdbwd_kernel(const float *__restrict__ P, const float *__restrict__ feat,
                                                        const float *__restrict__ target, const int *__restrict__ ok,
                                                        int count, int start, float inv_b, float *partial)
{
    __shared__ float w[MLP_NP];
    // per-sample rows, padded by one word so that a column walk is conflict-free
    __shared__ float sx[TB][MLP_IN + 1], sh1[TB][MLP_H + 1], sg1[TB][MLP_H + 1], sh2[TB][MLP_H + 1],
        sg2[TB][MLP_H + 1];
    __shared__ float sdy[TB], sloss[TB];
    const int t = threadIdx.x;
    for (int p = t; p < MLP_NP; p += TB) {
        const float v = P[p];
        const bool bias = (p >= MLP_OFF_B1 && p < MLP_OFF_W2) || (p >= MLP_OFF_B2 && p < MLP_OFF_W3) || p == MLP_OFF_B3;
        w[p] = bias ? v : bf16_value(v);
    }
    __syncthreads();

    const long s = (long)start + (long)blockIdx.x * TB + t;
    const bool valid = s < count && (ok == nullptr || ok[s] != 0);
    float x[MLP_IN], y = 0.f;
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) {
        const float v = valid ? feat[s * MLP_IN + k] : 0.f;
        const float hi = bf16_value(v);
        x[k] = hi + bf16_value(v - hi);
        sx[t][k] = x[k];
    }
    if (valid) y = target[s];
    // forward; sg1 / sg2 first hold GELU'(a1) / GELU'(a2)
    for (int o = 0; o < MLP_H; ++o) {
        float acc = w[MLP_OFF_B1 + o];
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) acc += w[MLP_OFF_W1 + o * MLP_IN + k] * x[k];
        sh1[t][o] = bf16_value(gelu_f(acc));
        sg1[t][o] = gelu_grad(acc);
    }
    float yhat = w[MLP_OFF_B3];
    for (int o = 0; o < MLP_H; ++o) {
        float acc = w[MLP_OFF_B2 + o];
#pragma unroll 8
        for (int i = 0; i < MLP_H; ++i) acc += w[MLP_OFF_W2 + o * MLP_H + i] * sh1[t][i];
        const float h2 = gelu_f(acc);
        sh2[t][o] = h2;
        sg2[t][o] = gelu_grad(acc);
        yhat += w[MLP_OFF_W3 + o] * h2;
    }
    // backward (the bf16 roundings pass gradients straight through)
    const float r = valid ? yhat - y : 0.f;
    const float dy = 2.0f * r * inv_b;
    for (int o = 0; o < MLP_H; ++o) sg2[t][o] = dy * w[MLP_OFF_W3 + o] * sg2[t][o];
    for (int i = 0; i < MLP_H; ++i) {
        float acc = 0.f;
#pragma unroll 8
        for (int o = 0; o < MLP_H; ++o) acc += w[MLP_OFF_W2 + o * MLP_H + i] * sg2[t][o];
        sg1[t][i] = acc * sg1[t][i];
    }
    sdy[t] = dy;
    sloss[t] = r * r;
    __syncthreads();

    float *out = partial + (size_t)blockIdx.x * PSTRIDE;
    for (int p = t; p < MLP_NP; p += TB) {
        float acc = 0.f;
        if (p < MLP_OFF_B1) {
            const int o = p / MLP_IN, k = p % MLP_IN;
            for (int q = 0; q < TB; ++q) acc += sg1[q][o] * sx[q][k];
        } else if (p < MLP_OFF_W2) {
            const int o = p - MLP_OFF_B1;
            for (int q = 0; q < TB; ++q) acc += sg1[q][o];
        } else if (p < MLP_OFF_B2) {
            const int o = (p - MLP_OFF_W2) / MLP_H, i = (p - MLP_OFF_W2) % MLP_H;
            for (int q = 0; q < TB; ++q) acc += sg2[q][o] * sh1[q][i];
        } else if (p < MLP_OFF_W3) {
            const int o = p - MLP_OFF_B2;
            for (int q = 0; q < TB; ++q) acc += sg2[q][o];
        } else if (p < MLP_OFF_B3) {
            const int o = p - MLP_OFF_W3;
            for (int q = 0; q < TB; ++q) acc += sdy[q] * sh2[q][o];
        } else {
            for (int q = 0; q < TB; ++q) acc += sdy[q];
        }
        out[p] = acc;
    }
    if (t == 0) {
        float l = 0.f;
        for (int q = 0; q < TB; ++q) l += sloss[q];
        out[MLP_NP] = l;
    }
}

// Sums the per-block partials in block order (deterministic).  mode 0: Adam
// step on P; mode 1: write the gradient to grad.  The loss sum is added to
// *loss_acc.
__global__ void mlp_adam_kernel(const float *__restrict__ partial, int nblocks, float *P, float *m1, float *m2,
                                float lr, float bc1, float bc2, int mode, float *grad, float *loss_acc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > MLP_NP) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[(size_t)b * PSTRIDE + p];
    if (p == MLP_NP) { *loss_acc += g; return; }
    if (mode == 1) { grad[p] = g; return; }
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float v = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = v;
    P[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// ---- inference and the guessed Riemann solve ---------------------------------------
__global__ void __launch_bounds__(256) mlp_infer_kernel(const MlpPacked *W, const float *feat, int count, float *out)
{
    __shared__ uint32_t lds[4][64 * 8];
    const int k = blockIdx.x * 256 + threadIdx.x;
    float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (k < count)
        for (int q = 0; q < MLP_IN; ++q) x[q] = feat[(size_t)k * MLP_IN + q];
    const float y = mlp_eval_wave(W, x, lds[threadIdx.x >> 6]);
    if (k < count) out[k] = y;
}

// in: rho_L u_L p_L rho_R u_R p_R gamma_L gamma_R eps tol per interface;
// out: u* p* crw0 crw1 iterations starting-guess (-1: the reference's guess);
// conv: the converge rule (riemann_exact)
__global__ void __launch_bounds__(256) riemann_guess_kernel(const double *in, int count, int N, int conv,
                                                            const MlpPacked *W, double *out)
{
    __shared__ uint32_t lds[4][64 * 8];
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool act = k < count;
    const double *a = in + (size_t)(act ? k : 0) * 10;
    const double rho_L = a[0], u_L = a[1], p_L = a[2], rho_R = a[3], u_R = a[4], p_R = a[5], gL = a[6], gR = a[7];
    const double eps = a[8], tol = a[9];
    double pg = -1.0;
    if (W) pg = learned_guess_wave(W, act, rho_L, u_L, p_L, rho_R, u_R, p_R, gL, gR, lds[threadIdx.x >> 6]);
    if (!act) return;
    const double c_L = sqrt(gL * p_L / rho_L), c_R = sqrt(gR * p_R / rho_R);
    const Gas GL = make_gas(gL), GR = make_gas(gR);
    const StarResult r = conv ? riemann_exact<true>(GL, GR, u_L, u_R, p_L, p_R, c_L, c_R, eps, tol, N, pg)
                              : riemann_exact<false>(GL, GR, u_L, u_R, p_L, p_R, c_L, c_R, eps, tol, N, pg);
    double *o = out + (size_t)k * 6;
    o[0] = r.u_star; o[1] = r.p_star; o[2] = r.crw0; o[3] = r.crw1; o[4] = r.iters; o[5] = pg;
}

// ---- C ABI -------------------------------------------------------------------------
namespace {
template <class T>
struct Dev {
    T *p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)); }
};
}  // namespace

extern "C" int hgrp_mlp_data(unsigned long long seed, int count, double tol, float *feat, float *target,
                             double *states, int *ok)
{
    if (count < 1 || !feat || !target || !ok || !(tol > 0.0)) return -1000;
    Dev<float> df, dt;
    Dev<double> ds;
    Dev<int> dk;
    MCHECK(df.alloc((size_t)count * MLP_IN));
    MCHECK(dt.alloc(count));
    MCHECK(dk.alloc(count));
    if (states) MCHECK(ds.alloc((size_t)count * 9));
    hipLaunchKernelGGL(mlp_data_kernel, dim3((count + 255) / 256), dim3(256), 0, 0, (uint64_t)seed, count, tol, df.p,
                       dt.p, states ? ds.p : nullptr, dk.p);
    MCHECK(hipGetLastError());
    MCHECK(hipDeviceSynchronize());
    MCHECK(hipMemcpy(feat, df.p, (size_t)count * MLP_IN * 4, hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(target, dt.p, (size_t)count * 4, hipMemcpyDeviceToHost));
    MCHECK(hipMemcpy(ok, dk.p, (size_t)count * 4, hipMemcpyDeviceToHost));
    if (states) MCHECK(hipMemcpy(states, ds.p, (size_t)count * 9 * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hgrp_mlp_loss_grad(const float *params, const float *feat, const float *target, int count,
                                  double *loss, float *grad)
{
    if (count < 1 || !params || !feat || !target || !grad) return -1000;
    const int nb = (count + TB - 1) / TB;
    Dev<float> dp, df, dt, part, dg, dl;
    MCHECK(dp.alloc(MLP_NP));
    MCHECK(df.alloc((size_t)count * MLP_IN));
    MCHECK(dt.alloc(count));
    MCHECK(part.alloc((size_t)nb * PSTRIDE));
    MCHECK(dg.alloc(MLP_NP));
    MCHECK(dl.alloc(1));
    MCHECK(hipMemcpy(dp.p, params, MLP_NP * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(df.p, feat, (size_t)count * MLP_IN * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(dt.p, target, (size_t)count * 4, hipMemcpyHostToDevice));
    MCHECK(hipMemset(dl.p, 0, 4));
    hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(nb), dim3(TB), 0, 0, dp.p, df.p, dt.p, (const int *)nullptr, count, 0,
                       1.0f / (float)count, part.p);
    MCHECK(hipGetLastError());
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NP + 1 + 255) / 256), dim3(256), 0, 0, part.p, nb, dp.p,
                       (float *)nullptr, (float *)nullptr, 0.f, 1.f, 1.f, 1, dg.p, dl.p);
    MCHECK(hipGetLastError());
    MCHECK(hipDeviceSynchronize());
    MCHECK(hipMemcpy(grad, dg.p, MLP_NP * 4, hipMemcpyDeviceToHost));
    float l = 0.f;
    MCHECK(hipMemcpy(&l, dl.p, 4, hipMemcpyDeviceToHost));
    if (loss) *loss = (double)l / (double)count;
    return 0;
}

extern "C" int hgrp_mlp_infer(const float *params, const float *feat, int count, float *out)
{
    if (count < 1 || !params || !feat || !out) return -1000;
    MlpPacked h;
    mlp_pack_host(params, &h);
    Dev<MlpPacked> dw;
    Dev<float> df, dout;
    MCHECK(dw.alloc(1));
    MCHECK(df.alloc((size_t)count * MLP_IN));
    MCHECK(dout.alloc(count));
    MCHECK(hipMemcpy(dw.p, &h, sizeof h, hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(df.p, feat, (size_t)count * MLP_IN * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((count + 255) / 256), dim3(256), 0, 0, dw.p, df.p, count, dout.p);
    MCHECK(hipGetLastError());
    MCHECK(hipDeviceSynchronize());
    MCHECK(hipMemcpy(out, dout.p, (size_t)count * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hgrp_riemann_guess_batch(const double *in, int count, int N, int rule, const float *params,
                                        double *out)
{
    if (count < 1 || !in || !out || N < 0 || (rule != HGRP_NEWTON_REFERENCE && rule != HGRP_NEWTON_CONVERGE))
        return -1000;
    Dev<MlpPacked> dw;
    if (params) {
        MlpPacked h;
        mlp_pack_host(params, &h);
        MCHECK(dw.alloc(1));
        MCHECK(hipMemcpy(dw.p, &h, sizeof h, hipMemcpyHostToDevice));
    }
    Dev<double> din, dout;
    MCHECK(din.alloc((size_t)count * 10));
    MCHECK(dout.alloc((size_t)count * 6));
    MCHECK(hipMemcpy(din.p, in, (size_t)count * 10 * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(riemann_guess_kernel, dim3((count + 255) / 256), dim3(256), 0, 0, din.p, count, N, rule,
                       params ? dw.p : (const MlpPacked *)nullptr, dout.p);
    MCHECK(hipGetLastError());
    MCHECK(hipDeviceSynchronize());
    MCHECK(hipMemcpy(out, dout.p, (size_t)count * 6 * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int hgrp_mlp_train(unsigned long long seed, int n_samples, int epochs, int batch, double lr, double tol,
                              float *params, float *epoch_loss)
{
    if (n_samples < batch || epochs < 1 || batch < TB || batch % TB || !params || !(tol > 0.0) || !(lr > 0.0))
        return -1000;
    const int nb = batch / TB;
    const int steps_per_epoch = n_samples / batch;
    Dev<float> df, dt, dp, m1, m2, part, dl;
    Dev<int> dk;
    MCHECK(df.alloc((size_t)n_samples * MLP_IN));
    MCHECK(dt.alloc(n_samples));
    MCHECK(dk.alloc(n_samples));
    MCHECK(dp.alloc(MLP_NP));
    MCHECK(m1.alloc(MLP_NP));
    MCHECK(m2.alloc(MLP_NP));
    MCHECK(part.alloc((size_t)nb * PSTRIDE));
    MCHECK(dl.alloc(epochs));
    hipStream_t st;
    MCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    MCHECK(hipMemcpyAsync(dp.p, params, MLP_NP * 4, hipMemcpyHostToDevice, st));
    MCHECK(hipMemsetAsync(m1.p, 0, MLP_NP * 4, st));
    MCHECK(hipMemsetAsync(m2.p, 0, MLP_NP * 4, st));
    MCHECK(hipMemsetAsync(dl.p, 0, (size_t)epochs * 4, st));
    hipLaunchKernelGGL(mlp_data_kernel, dim3((n_samples + 255) / 256), dim3(256), 0, st, (uint64_t)seed, n_samples,
                       tol, df.p, dt.p, (double *)nullptr, dk.p);
    MCHECK(hipGetLastError());
    std::vector<int> okh(n_samples);
    MCHECK(hipMemcpyAsync(okh.data(), dk.p, (size_t)n_samples * 4, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    long used = 0;
    for (int e = 0; e < epochs; ++e)
        for (int s = 0; s < steps_per_epoch; ++s)
            for (int q = 0; q < batch; ++q) used += okh[(size_t)s * batch + q] != 0;
    const long total = (long)epochs * steps_per_epoch;
    const double lr_min = lr * 0.01;
    long t = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int s = 0; s < steps_per_epoch; ++s, ++t) {
            const double lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + cos(M_PI * (double)t / (double)total));
            const float bc1 = (float)(1.0 - pow(0.9, (double)(t + 1)));
            const float bc2 = (float)(1.0 - pow(0.999, (double)(t + 1)));
            hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(nb), dim3(TB), 0, st, dp.p, df.p, dt.p, dk.p, n_samples,
                               s * batch, 1.0f / (float)batch, part.p);
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NP + 1 + 255) / 256), dim3(256), 0, st, part.p, nb, dp.p,
                               m1.p, m2.p, (float)lr_t, bc1, bc2, 0, (float *)nullptr, dl.p + e);
        }
        MCHECK(hipGetLastError());
    }
    MCHECK(hipMemcpyAsync(params, dp.p, MLP_NP * 4, hipMemcpyDeviceToHost, st));
    std::vector<float> lh(epochs);
    MCHECK(hipMemcpyAsync(lh.data(), dl.p, (size_t)epochs * 4, hipMemcpyDeviceToHost, st));
    MCHECK(hipStreamSynchronize(st));
    if (epoch_loss)
        for (int e = 0; e < epochs; ++e) epoch_loss[e] = (float)(lh[e] * (double)epochs / (double)(used ? used : 1));
    return 0;
}
