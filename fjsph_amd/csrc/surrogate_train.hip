/*
 * Training of the MLP surrogate on the GPU with the project's own kernels
 * (it has no counterpart in the reference).  One step over a batch of B
 * samples (B a multiple of 64):
 *
 *   k_mlp_batch    sample indices of the step (counter-based hash of seed,
 *                  step, position)
 *   k_mlp_train    one wave per 64 samples: forward pass as in inference
 *                  (BF16 weights and activations, FP32 accumulation),
 *                  squared-error loss, backward pass with the same MFMA
 *                  product, per-wave partial gradients
 *   k_mlp_reduce   gradients = sum of the per-wave partials, in wave order
 *                  (deterministic)
 *   k_mlp_adam     Adam on the FP32 master parameters
 *   k_mlp_pack     BF16 A-layout tiles of the weights (forward) and of
 *                  their transposes (backward) from the masters
 *
 * Backward of layer l (Z_l = W_l H_{l-1} + b_l, H_l = gelu(Z_l); the
 * output layer Y = W_o H_L + b_o has no activation), loss = sum (Y - T)^2
 * over the batch and outputs, dY = 2 (Y - T) / (B (D + 2)):
 *   dZ_l = dH_l * gelu'(Z_l)          FP32, rounded to BF16 for the products
 *   db_l = sum over points of dZ_l    FP32
 *   dW_l = dZ_l H_{l-1}^T             BF16 x BF16, FP32 accumulation
 *   dH_{l-1} = W_l^T dZ_l             BF16 x BF16, FP32 accumulation
 * Z_l (FP32) and H_l (BF16) of the forward pass are kept in global memory
 * for the backward pass.  The loss reported is the mean squared error
 * over the batch and outputs.
 */
#include "mlp_common.hpp"

namespace fj {

/* Counter-based 64-bit hash (splitmix64 finaliser). */
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(TPB) void k_mlp_batch(uint64_t seed, int64_t step, int64_t n, int32_t batch,
                                                   int32_t* __restrict__ idx)
{
    const int32_t j = blockIdx.x * TPB + threadIdx.x;
    if (j >= batch) return;
    const uint64_t h = mix64(seed ^ mix64(uint64_t(step) * 0x100000001b3ull + uint64_t(j)));
    idx[j] = int32_t(h % uint64_t(n));
}

/* B rows of the current layer's input, feature-major: rows[f * SROW + pt]
 * from a point-major global array g[(base + pt) * 64 + f], f < nf. */
__device__ __forceinline__ void load_feature_major(const __bf16* __restrict__ g, int64_t base, int nf, __bf16* rows,
                                                   int lane)
{
    for (int f = 0; f < nf; ++f) rows[f * SROW + lane] = g[(base + lane) * SW + f];
    for (int f = nf; f < 16 * ((nf + 15) / 16); ++f) rows[f * SROW + lane] = __bf16(0.f);
}

/* acc[MT][NT] (D layout: features x points or outputs x inputs) =
 * A rows (feature-major, K = 64 points) x B rows (feature-major, K = 64
 * points): C[i][j] = sum_pt A[i][pt] B[j][pt]. */
template <int MT, int NT>
__device__ __forceinline__ void outer_product(const __bf16* a_rows, const __bf16* b_rows, f32x4 (&acc)[4][4], int lane)
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) {
        bf16x8 b[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            b[nt] = *(const bf16x8*)(b_rows + (nt * 16 + (lane & 15)) * SROW + kc * 32 + 8 * (lane >> 4));
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const bf16x8 a = *(const bf16x8*)(a_rows + (mt * 16 + (lane & 15)) * SROW + kc * 32 + 8 * (lane >> 4));
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mt][nt], 0, 0, 0);
        }
    }
}

/* Write the per-wave weight-gradient partial of a layer (nout x nin, row
 * major) from the D layout of outer_product. */
template <int MT, int NT>
__device__ __forceinline__ void store_dw(const f32x4 (&acc)[4][4], int nout, int nin, float* __restrict__ out,
                                         int lane)
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int fo = d_feature(mt, lane, e), fi = d_point(nt, lane);
                if (fo < nout && fi < nin) out[fo * nin + fi] = acc[mt][nt][e];
            }
}

/* Sum over the 64 points of each feature row of a D-layout tile set
 * (features f = d_feature(mt, lane, e)); lanes with lane % 16 == 0 write
 * out[f] for f < nf. */
template <int MT>
__device__ __forceinline__ void store_db(const float (&v)[4][4][4], int nf, float* __restrict__ out, int lane)
{
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float sum = v[mt][0][e] + v[mt][1][e] + v[mt][2][e] + v[mt][3][e];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) sum += __shfl_xor(sum, o, 64);
            const int f = d_feature(mt, lane, e);
            if ((lane & 15) == 0 && f < nf) out[f] = sum;
        }
}

/* One training pass over 64 samples per block of one wave.  Parameter
 * offsets (row major, as fjsph_set_surrogate takes them): W_1 [64][D],
 * W_2..W_L [64][64], W_o [D+2][64]; b_1..b_L [64], b_o [D+2]. */
template <int D>
__global__ __launch_bounds__(64) void k_mlp_train(MlpTrain m)
{
    constexpr int NO = D + 2;
    __shared__ __attribute__((aligned(16))) __bf16 pm[64 * SROW];  /* point-major rows   */
    __shared__ __attribute__((aligned(16))) __bf16 fa[64 * SROW];  /* feature-major: dZ  */
    __shared__ __attribute__((aligned(16))) __bf16 fb[64 * SROW];  /* feature-major: H   */
    __shared__ int32_t sidx[64];
    const int lane = threadIdx.x;
    const int L = m.n_hidden;
    const int64_t w = blockIdx.x;
    const int64_t base = w * 64;
    const int32_t my = m.idx[base + lane];
    sidx[lane] = my;
    const bf16x8* Wf = (const bf16x8*)m.wf;
    const bf16x8* Wt = (const bf16x8*)m.wt;
    const int64_t B = m.batch;
    /* ---- forward */
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        const __bf16 v = __bf16(k < D ? m.X[(int64_t)my * 4 + k] : 0.f);
        pm[lane * SROW + k] = v;
        if (k < SW) m.H[(base + lane) * SW + k] = v;
    }
    __builtin_amdgcn_wave_barrier();
    f32x4 acc[4][4];
    for (int l = 1; l <= L; ++l) {
        if (l == 1) mlp_layer<4, 1>(Wf, pm, acc, lane);
        else mlp_layer<4, 2>(Wf + (size_t)(4 + (l - 2) * 8) * 64, pm, acc, lane);
        __builtin_amdgcn_wave_barrier();
        float* Z = m.Z + (size_t)(l - 1) * B * SW;
        __bf16* H = m.H + (size_t)l * B * SW;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int pt = d_point(nt, lane);
                float4 z;
                __bf16 hv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int f = d_feature(mt, lane, e);
                    const float zz = acc[mt][nt][e] + m.b[(l - 1) * SW + f];
                    (&z.x)[e] = zz;
                    hv[e] = __bf16(gelu_erf(zz));
                    pm[pt * SROW + f] = hv[e];
                }
                const int f0 = d_feature(mt, lane, 0);
                *(float4*)(Z + (base + pt) * SW + f0) = z;
#pragma unroll
                for (int e = 0; e < 4; ++e) H[(base + pt) * SW + f0 + e] = hv[e];
            }
        __builtin_amdgcn_wave_barrier();
    }
    mlp_layer<1, 2>(Wf + (size_t)(4 + (L - 1) * 8) * 64, pm, acc, lane);
    /* the backward pass reads rows of Z and H that other lanes stored:
     * complete the stores and drop stale L1 lines */
    __threadfence();
    /* ---- loss and dY (outputs f = d_feature(0, lane, e) < NO) */
    float dy[4][4][4];
    float lsum = 0.f;
    const float gscale = 2.0f / (float(B) * float(NO));
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int f = d_feature(0, lane, e), pt = d_point(nt, lane);
            float g = 0.f;
            if (f < NO) {
                const float diff = acc[0][nt][e] + m.bo[f] - m.T[(int64_t)sidx[pt] * 8 + f];
                lsum += diff * diff;
                g = gscale * diff;
            }
            dy[0][nt][e] = g;
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0) m.lossp[w] = lsum;
    float* dWp = m.dWp + (size_t)w * m.nw;
    float* dbp = m.dbp + (size_t)w * m.nb;
    store_db<1>(dy, NO, dbp + (size_t)L * SW, lane);
    /* dY to LDS as BF16: point-major rows (K = outputs, padded to 32) and
     * feature-major rows (outputs x points, padded to 16 rows) */
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 32; ++k) pm[lane * SROW + k] = __bf16(0.f);
#pragma unroll
    for (int k = 0; k < 16; ++k) fa[k * SROW + lane] = __bf16(0.f);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int f = d_feature(0, lane, e), pt = d_point(nt, lane);
            const __bf16 v = __bf16(dy[0][nt][e]);
            pm[pt * SROW + f] = v;
            fa[f * SROW + pt] = v;
        }
    load_feature_major(m.H + (size_t)L * B * SW, base, SW, fb, lane);
    __builtin_amdgcn_wave_barrier();
    /* output layer: dW_o = dY H_L^T, dH_L = W_o^T dY */
    {
        f32x4 g[4][4];
        outer_product<1, 4>(fa, fb, g, lane);
        store_dw<1, 4>(g, NO, SW, dWp + (size_t)(SW * D + (L - 1) * SW * SW), lane);
    }
    mlp_layer<4, 1>(Wt + (size_t)((L - 1) * 8) * 64, pm, acc, lane);
    /* hidden layers, last to first */
    for (int l = L; l >= 1; --l) {
        __builtin_amdgcn_wave_barrier();
        const float* Z = m.Z + (size_t)(l - 1) * B * SW;
        float dz[4][4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int pt = d_point(nt, lane), f0 = d_feature(mt, lane, 0);
                const float4 z = *(const float4*)(Z + (base + pt) * SW + f0);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = acc[mt][nt][e] * gelu_erf_grad((&z.x)[e]);
                    dz[mt][nt][e] = v;
                    const __bf16 vb = __bf16(v);
                    pm[pt * SROW + f0 + e] = vb;
                    fa[(f0 + e) * SROW + pt] = vb;
                }
            }
        store_db<4>(dz, SW, dbp + (size_t)(l - 1) * SW, lane);
        const int nin = l == 1 ? D : SW;
        load_feature_major(m.H + (size_t)(l - 1) * B * SW, base, nin, fb, lane);
        __builtin_amdgcn_wave_barrier();
        f32x4 g[4][4];
        float* dW = dWp + (l == 1 ? 0 : (size_t)(SW * D + (l - 2) * SW * SW));
        if (l == 1) {
            outer_product<4, 1>(fa, fb, g, lane);
            store_dw<4, 1>(g, SW, D, dW, lane);
        } else {
            outer_product<4, 4>(fa, fb, g, lane);
            store_dw<4, 4>(g, SW, SW, dW, lane);
            mlp_layer<4, 2>(Wt + (size_t)((l - 2) * 8) * 64, pm, acc, lane);
        }
    }
}

/* grad[k] = sum over the waves of part[w][k], in wave order. */
__global__ __launch_bounds__(TPB) void k_mlp_reduce(const float* __restrict__ part, int64_t nwaves, int64_t stride,
                                                    int64_t n, float* __restrict__ grad)
{
    const int64_t k = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (k >= n) return;
    float s = 0.f;
    for (int64_t w = 0; w < nwaves; ++w) s += part[w * stride + k];
    grad[k] = s;
}

/* Adam (Kingma and Ba) on p[0..n): bias-corrected moments, step t >= 1. */
__global__ __launch_bounds__(TPB) void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                                                  float* __restrict__ m1, float* __restrict__ m2, int64_t n, float lr,
                                                  float b1, float b2, float eps, float c1, float c2)
{
    const int64_t k = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk;
    const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    p[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

/* BF16 A-layout tiles of the weights: forward tiles (layer 1: 4 tiles of
 * W_1 padded to 64 x 32; hidden: 8 tiles each; output: 2 tiles of W_o
 * padded to 16 x 64), then backward tiles of the transposes (hidden l >= 2:
 * 8 tiles of W_l^T; output: 4 tiles of W_o^T padded to 64 x 32).  One
 * thread per (tile, lane). */
template <int D>
__global__ __launch_bounds__(TPB) void k_mlp_pack(const float* __restrict__ w, int32_t n_hidden, bf16x8* __restrict__ wf,
                                                  bf16x8* __restrict__ wt)
{
    constexpr int NO = D + 2;
    const int L = n_hidden;
    const int nf = 4 + (L - 1) * 8 + 2, nb = (L - 1) * 8 + 4;
    const int64_t gid = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (gid >= int64_t(nf + nb) * 64) return;
    const int tile = int(gid / 64), lane = int(gid % 64);
    const int i = lane & 15, k0 = 8 * (lane >> 4);
    bf16x8 v;
    if (tile < nf) {
        /* forward: A[i][k] = W[mt*16 + i][kc*32 + k] */
        int mt, kc, nout, nin;
        size_t off;
        if (tile < 4) {
            mt = tile; kc = 0; nout = SW; nin = D; off = 0;
        } else if (tile < 4 + (L - 1) * 8) {
            const int l = (tile - 4) / 8, t = (tile - 4) % 8;
            mt = t / 2; kc = t % 2; nout = SW; nin = SW; off = size_t(SW) * D + size_t(l) * SW * SW;
        } else {
            mt = 0; kc = tile - 4 - (L - 1) * 8; nout = NO; nin = SW; off = size_t(SW) * D + size_t(L - 1) * SW * SW;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int r = mt * 16 + i, c = kc * 32 + k0 + e;
            v[e] = __bf16(r < nout && c < nin ? w[off + size_t(r) * nin + c] : 0.f);
        }
        wf[gid] = v;
    } else {
        /* backward: A[i][k] = W^T[mt*16 + i][kc*32 + k] = W[kc*32 + k][mt*16 + i] */
        const int tb = tile - nf;
        int mt, kc, nout, nin;
        size_t off;
        if (tb < (L - 1) * 8) {
            const int l = tb / 8, t = tb % 8;
            mt = t / 2; kc = t % 2; nout = SW; nin = SW; off = size_t(SW) * D + size_t(l) * SW * SW;
        } else {
            mt = tb - (L - 1) * 8; kc = 0; nout = NO; nin = SW; off = size_t(SW) * D + size_t(L - 1) * SW * SW;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int r = kc * 32 + k0 + e, c = mt * 16 + i; /* W[r][c], r < nout, c < nin */
            v[e] = __bf16(r < nout && c < nin ? w[off + size_t(r) * nin + c] : 0.f);
        }
        wt[gid - int64_t(nf) * 64] = v;
    }
}

void launch_mlp_batch(uint64_t seed, int64_t step, int64_t n, int32_t batch, int32_t* idx, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_batch, dim3(grid_of(batch)), dim3(TPB), 0, st, seed, step, n, batch, idx);
}

void launch_mlp_train(int dim, const MlpTrain& m, hipStream_t st)
{
    const dim3 g((unsigned)(m.batch / 64)), b(64);
    if (dim == 2) hipLaunchKernelGGL(k_mlp_train<2>, g, b, 0, st, m);
    else hipLaunchKernelGGL(k_mlp_train<3>, g, b, 0, st, m);
}

void launch_mlp_reduce(const float* part, int64_t nwaves, int64_t stride, int64_t n, float* grad, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_reduce, dim3(grid_of(n)), dim3(TPB), 0, st, part, nwaves, stride, n, grad);
}

void launch_mlp_adam(float* p, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid_of(n)), dim3(TPB), 0, st, p, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
}

int64_t mlp_tiles_fwd(int n_hidden) { return 4 + int64_t(n_hidden - 1) * 8 + 2; }
int64_t mlp_tiles_bwd(int n_hidden) { return int64_t(n_hidden - 1) * 8 + 4; }

void launch_mlp_pack(int dim, const float* w, int32_t n_hidden, void* wf, void* wt, hipStream_t st)
{
    const int64_t n = (mlp_tiles_fwd(n_hidden) + mlp_tiles_bwd(n_hidden)) * 64;
    if (dim == 2)
        hipLaunchKernelGGL(k_mlp_pack<2>, dim3(grid_of(n)), dim3(TPB), 0, st, w, n_hidden, (bf16x8*)wf, (bf16x8*)wt);
    else
        hipLaunchKernelGGL(k_mlp_pack<3>, dim3(grid_of(n)), dim3(TPB), 0, st, w, n_hidden, (bf16x8*)wf, (bf16x8*)wt);
}

}  // namespace fj
