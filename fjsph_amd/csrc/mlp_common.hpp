/*
 * Building blocks of the MLP surrogate kernels (inference and training):
 * one wave computes a 64 x 64 (features x points) layer from 16 x 16 output
 * tiles with v_mfma_f32_16x16x32_bf16, BF16 operands and FP32 accumulation.
 * Operand layout of that instruction (lane l = 0..63):
 *     A[i][k]  (16 x 32): lane i + 16*(k/8), element k%8
 *     B[k][j]  (32 x 16): lane j + 16*(k/8), element k%8
 *     D[i][j]  (16 x 16): lane j + 16*(i/4), element i%4
 * Weights come in A-layout tiles (8 BF16 per lane, one 16-byte load per
 * tile); activations sit in LDS as BF16 rows of SROW elements, one row per
 * point ("point-major", the B operand of a layer) or per feature
 * ("feature-major", the operands of the weight-gradient products).
 */
#pragma once
#include "device_common.hpp"

namespace fj {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SW = 64;       /* hidden width */
constexpr int SROW = SW + 8; /* LDS row, padded against bank conflicts */

/* erf by Abramowitz and Stegun 7.1.26 (|error| <= 1.5e-7): for x >= 0,
 * erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2),
 * t = 1 / (1 + p x); odd in x.  Far below the BF16 rounding of the
 * activations it feeds. */
__device__ __forceinline__ float erf_as(float x)
{
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * ax);
    const float poly =
        t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
    const float y = 1.0f - poly * __expf(-ax * ax);
    return copysignf(y, x);
}

__device__ __forceinline__ float gelu_erf(float x) { return 0.5f * x * (1.0f + erf_as(x * 0.70710678118654752f)); }

/* d gelu / dx = Phi(x) + x phi(x) */
__device__ __forceinline__ float gelu_erf_grad(float x)
{
    return 0.5f * (1.0f + erf_as(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

/* out[MT tiles of 16 features][4 tiles of 16 points] = W (A tiles, KC
 * chunks of 32 inputs) x rows (point-major LDS rows, B operand) */
template <int MT, int KC>
__device__ __forceinline__ void mlp_layer(const bf16x8* __restrict__ wA, const __bf16* rows, f32x4 (&acc)[4][4],
                                          int lane)
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
        bf16x8 b[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            b[nt] = *(const bf16x8*)(rows + (nt * 16 + (lane & 15)) * SROW + kc * 32 + 8 * (lane >> 4));
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const bf16x8 a = wA[(mt * KC + kc) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
        }
    }
}

/* Feature of element e of tile (mt, nt) in the D layout, and its point. */
__device__ __forceinline__ int d_feature(int mt, int lane, int e) { return mt * 16 + 4 * (lane >> 4) + e; }
__device__ __forceinline__ int d_point(int nt, int lane) { return nt * 16 + (lane & 15); }

}  // namespace fj
