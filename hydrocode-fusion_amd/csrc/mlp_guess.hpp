// mlp_guess.hpp -- learned starting guess for the Newton solve of p*.
//
// No counterpart in the reference: its exact solver (Riemann_solver_exact,
// src/riemann_solver/riemann_solver_exact_Ben.c:87-93) always starts Newton
// from half the intersection of the two acoustic tangents.  Here a small MLP
// predicts y = log(p* / p_TR), the correction to the two-rarefaction estimate
// p_TR (exact when both waves are rarefactions of one gas; two Newton steps
// from the mean-exponent form for two gases), from six dimensionless inputs
//   log(rho_L/rho_R)/ln 1e3, log(p_L/p_R)/ln 1e4, (u_R-u_L)/(5 c_L),
//   log(c_R/c_L)/5, gamma_L - 1.4, gamma_R - 1.4,
// and Newton starts from p_TR * exp(y).  The exact solver decides p*: after at
// most GUESS_ITERS iterations from the learned guess without meeting the
// tolerance it restarts from the reference's guess.
//
// Network: 6 -> 32 -> 32 -> 1, GELU (tanh form) on the hidden layers, BF16
// weights, FP32 accumulation.  A wave evaluates it for its 64 lanes with
// v_mfma_f32_16x16x16_bf16 in the transposed form H^T = W . X^T, so that the
// accumulator fragment of one layer (lane l: rows 4(l/16)+j, sample l%16) is
// exactly the B fragment the next layer needs: no transposes between layers.
// The inputs enter as a BF16 hi/lo pair (K slots 0-5 and 6-11 of the first
// layer), which keeps 16 bits of them; biases enter as the FP32 accumulator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hgrp {

constexpr int MLP_IN = 6, MLP_H = 32;
// master parameter vector: W1[32][6] b1[32] W2[32][32] b2[32] w3[32] b3
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 192, MLP_OFF_W2 = 224, MLP_OFF_B2 = 1248, MLP_OFF_W3 = 1280,
              MLP_OFF_B3 = 1312, MLP_NP = 1313;

// Packed inference weights (built from the master vector by mlp_pack_host):
// per-lane MFMA A fragments of W1 (2 row blocks) and W2 (2 row x 2 column
// blocks), 4 bf16 each; FP32 biases and bf16-rounded w3.
struct MlpPacked {
    uint2 w1f[2][64];
    uint2 w2f[2][2][64];
    float b1[MLP_H], b2[MLP_H], w3[MLP_H], b3;
};

// master parameter vector -> packed inference weights (host side)
void mlp_pack_host(const float *P, MlpPacked *W);

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint32_t bf16_bits(float f)
{
    uint32_t b;
    __builtin_memcpy(&b, &f, 4);
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}
__host__ __device__ inline float bf16_value(float f)
{
    uint32_t b = bf16_bits(f) << 16;
    float r;
    __builtin_memcpy(&r, &b, 4);
    return r;
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// two floats -> their bf16 bits packed (lo in bits 0-15), round to nearest
// even: one v_cvt_pk_bf16_f32 on gfx950 (bf16_bits' rounding for finite values)
__device__ __forceinline__ uint32_t bf16_pack2(float lo, float hi)
{
    const f32x2_t v = {lo, hi};
    const bf16x2_t b = __builtin_convertvector(v, bf16x2_t);
    uint32_t r;
    __builtin_memcpy(&r, &b, 4);
    return r;
}

// GELU, tanh form: x * sigmoid(2a), a = sqrt(2/pi) (x + 0.044715 x^3); the
// sigmoid from the hardware reciprocal (v_rcp_f32, within 1 ulp) instead of an
// IEEE division, which costs about ten instructions: 64 of these run per lane
// in each wave that evaluates the network
__device__ __forceinline__ float gelu_f(float x)
{
    const float a = 0.7978845608f * (x + 0.044715f * x * x * x);
    return x * __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * a));
}

// Inputs of the network and the two-rarefaction estimate, from the states
// normal to the interface.  p_TR solves
//   2c_L/(g_L-1) ((p/p_L)^z_L - 1) + 2c_R/(g_R-1) ((p/p_R)^z_R - 1) + (u_R - u_L) = 0,
// z = (g-1)/(2g): the closed form with the mean exponent (exact for one gas),
// then two Newton steps in log p for two gases.  Returns false where no
// estimate exists (the data would open a vacuum).
__device__ __forceinline__ bool mlp_features(double rho_L, double u_L, double p_L, double rho_R, double u_R,
                                             double p_R, double g_L, double g_R, float x[MLP_IN], float *p_tr)
{
    const float rl = (float)rho_L, rr = (float)rho_R, pl = (float)p_L, pr = (float)p_R;
    const float gl = (float)g_L, gr = (float)g_R;
    const float cl = sqrtf(gl * pl / rl), cr = sqrtf(gr * pr / rr);
    const float du = (float)(u_R - u_L);
    const float lrl = __logf(rl), lrr = __logf(rr), lpl = __logf(pl), lpr = __logf(pr);
    x[0] = (lrl - lrr) * 0.1447648273f;      // 1 / ln 1e3
    x[1] = (lpl - lpr) * 0.1085736205f;      // 1 / ln 1e4
    x[2] = du / cl * 0.2f;
    x[3] = 0.5f * ((lpr - lrr + __logf(gr)) - (lpl - lrl + __logf(gl))) * 0.2f;
    x[4] = gl - 1.4f;
    x[5] = gr - 1.4f;
    const float zl = (gl - 1.0f) / (2.0f * gl), zr = (gr - 1.0f) / (2.0f * gr), z = 0.5f * (zl + zr);
    const float al = 2.0f * cl / (gl - 1.0f), ar = 2.0f * cr / (gr - 1.0f);
    const float num = al + ar - du;
    const float den = al * __expf(-z * lpl) + ar * __expf(-z * lpr);
    bool ok = num > 0.0f && den > 0.0f && cl > 0.0f && cr > 0.0f && __builtin_isfinite(x[0] + x[1] + x[2] + x[3]);
    float y = ok ? __logf(num / den) / z : 0.0f;             // log p
    if (ok && gl != gr) {
        for (int it = 0; it < 2; ++it) {
            const float ql = __expf(zl * (y - lpl)), qr = __expf(zr * (y - lpr));
            const float F = al * (ql - 1.0f) + ar * (qr - 1.0f) + du;
            const float dF = al * zl * ql + ar * zr * qr;
            y -= F / dF;
        }
    }
    *p_tr = ok ? __expf(y) : 0.0f;
    return ok && *p_tr > 0.0f && __builtin_isfinite(*p_tr);
}

__device__ __forceinline__ uint32_t pack2(uint32_t lo, uint32_t hi) { return lo | (hi << 16); }

// Evaluates the network for the 64 lanes of the calling wave; every lane must
// call it (uniform control flow).  x: this lane's inputs (zeros for lanes
// without work).  lds: 64 * 8 uint32 of LDS private to this wave.
__device__ inline float mlp_eval_wave(const MlpPacked *__restrict__ W, const float x[MLP_IN], uint32_t *lds)
{
    const int lane = threadIdx.x & 63;
    // this lane's sample as 16 bf16 K slots: hi parts, lo parts, zeros
    uint32_t hi[MLP_IN], lo[MLP_IN];
#pragma unroll
    for (int k = 0; k < MLP_IN; ++k) {
        hi[k] = bf16_bits(x[k]);
        uint32_t hb = hi[k] << 16;
        float hv;
        __builtin_memcpy(&hv, &hb, 4);
        lo[k] = bf16_bits(x[k] - hv);
    }
    uint32_t *mine = lds + lane * 8;
    mine[0] = pack2(hi[0], hi[1]); mine[1] = pack2(hi[2], hi[3]); mine[2] = pack2(hi[4], hi[5]);
    mine[3] = pack2(lo[0], lo[1]); mine[4] = pack2(lo[2], lo[3]); mine[5] = pack2(lo[4], lo[5]);
    mine[6] = 0u; mine[7] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const int g = lane >> 4, col = lane & 15;
    const uint2 a10 = W->w1f[0][lane], a11 = W->w1f[1][lane];
    float y_mine = 0.0f;
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
        // layer 1: B fragment = slots 4g..4g+3 of sample 16 sb + col
        const uint2 xb = *(const uint2 *)(lds + (16 * sb + col) * 8 + 2 * g);
        bf16x4 bx;
        __builtin_memcpy(&bx, &xb, 8);
        f32x4 h[2];
#pragma unroll
        for (int ob = 0; ob < 2; ++ob) {
            f32x4 c = {W->b1[16 * ob + 4 * g + 0], W->b1[16 * ob + 4 * g + 1], W->b1[16 * ob + 4 * g + 2],
                       W->b1[16 * ob + 4 * g + 3]};
            const uint2 af = ob ? a11 : a10;
            bf16x4 av;
            __builtin_memcpy(&av, &af, 8);
            h[ob] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bx, c, 0, 0, 0);
        }
        // GELU, round to bf16: the B fragments of layer 2 (K block = row block of layer 1)
        bf16x4 hb[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            uint2 u;
            u.x = bf16_pack2(gelu_f(h[kb][0]), gelu_f(h[kb][1]));
            u.y = bf16_pack2(gelu_f(h[kb][2]), gelu_f(h[kb][3]));
            __builtin_memcpy(&hb[kb], &u, 8);
        }
        float part = 0.0f;
#pragma unroll
        for (int ob = 0; ob < 2; ++ob) {
            f32x4 c = {W->b2[16 * ob + 4 * g + 0], W->b2[16 * ob + 4 * g + 1], W->b2[16 * ob + 4 * g + 2],
                       W->b2[16 * ob + 4 * g + 3]};
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                const uint2 af = W->w2f[ob][kb][lane];
                bf16x4 av;
                __builtin_memcpy(&av, &af, 8);
                c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, hb[kb], c, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) part += W->w3[16 * ob + 4 * g + j] * gelu_f(c[j]);
        }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        if (sb == g) y_mine = part + W->b3;
    }
    __builtin_amdgcn_wave_barrier();   // lds may be reused by the caller after this
    return y_mine;
}

// Learned starting pressure for this lane's interface, or -1 where the lane
// has none (want == false, or no two-rarefaction estimate).  Wave-uniform call.
// Weak interfaces (pressure, velocity and density jumps all within about 10 %)
// start from p_TR itself, which is already within O(jump^3) of p* and where
// the network's correction is nil, and so do two rarefactions of one gas; the
// network runs only in waves that hold another (strong) interface.
__device__ inline double learned_guess_wave(const MlpPacked *__restrict__ W, bool want, double rho_L, double u_L,
                                            double p_L, double rho_R, double u_R, double p_R, double g_L,
                                            double g_R, uint32_t *lds)
{
    float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float p_tr = 0.f;
    bool ok = false;
    if (want) ok = mlp_features(rho_L, u_L, p_L, rho_R, u_R, p_R, g_L, g_R, x, &p_tr);
    // |log(p_L/p_R)| < 0.1, |u_R-u_L| < 0.1 c_L, |log(rho_L/rho_R)| < 0.1, one gas
    // Two rarefactions of one gas (p_TR at or below both pressures): p_TR is then
    // p* itself, to its single-precision rounding, and the network's correction
    // would only add its fitting error
    const bool two_rare = ok && x[4] == x[5] && p_tr <= fminf((float)p_L, (float)p_R);
    const bool strong = ok && !two_rare &&
                        (fabsf(x[1]) > 0.0109f || fabsf(x[2]) > 0.02f || fabsf(x[0]) > 0.0145f || x[4] != x[5]);
    if (!strong) {
#pragma unroll
        for (int k = 0; k < MLP_IN; ++k) x[k] = 0.f;
    }
    float y = 0.f;
    if (__any(strong)) y = mlp_eval_wave(W, x, lds);
    if (!ok) return -1.0;
    const double p0 = (double)p_tr * (strong ? (double)__expf(y) : 1.0);
    return (p0 > 0.0 && p0 < 1e300) ? p0 : -1.0;
}

}  // namespace hgrp
