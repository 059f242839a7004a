/*
 * Neural surrogate of the aerodynamic mesh field (aero_source =
 * FJSPH_SURROGATE; it has no counterpart in the reference): an MLP
 *     h_1 = gelu(W_1 xn + b_1), h_k = gelu(W_k h_{k-1} + b_k), k = 2..L,
 *     y = W_o h_L + b_o,
 * width 64, GELU in its erf form, xn = (x - in_shift) / in_scale, and the
 * outputs (gas velocity, pressure, density) = y * out_scale + out_shift.
 *
 * Each wave evaluates 64 points, one layer at a time with the MFMA layer
 * product of mlp_common.hpp; the weights are stored on the host already in
 * the A layout, the activations sit in LDS as BF16 rows per point.
 */
#include "mlp_common.hpp"

namespace fj {

/* bias + GELU, back to the LDS rows (after every lane has read them); the
 * 4 elements of a lane's output tile are 4 consecutive features of one
 * point, one 8-byte store */
__device__ __forceinline__ void store_hidden(const f32x4 (&acc)[4][4], const float* __restrict__ bias, __bf16* act,
                                             int lane)
{
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int f0 = d_feature(mt, lane, 0);
        const float4 bb = *(const float4*)(bias + f0);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x4 v = {__bf16(gelu_erf(acc[mt][nt][0] + bb.x)), __bf16(gelu_erf(acc[mt][nt][1] + bb.y)),
                              __bf16(gelu_erf(acc[mt][nt][2] + bb.z)), __bf16(gelu_erf(acc[mt][nt][3] + bb.w))};
            *(bf16x4*)(act + d_point(nt, lane) * SROW + f0) = v;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

/* The aero particles of state s (FREE with lam_nb <= lam_cutoff) into a
 * segmented list: the slots of segment g = [g*AERO_SEG, (g+1)*AERO_SEG)
 * go to list[g*AERO_SEG .. g*AERO_SEG + counts[g*AERO_CSTRIDE]) (in no
 * particular order within the segment; each is looked up on its own), one
 * counter per segment on its own 128-byte line so the waves' appends do
 * not all meet at one address.  The others get no cell, as FindCell
 * does. */
__global__ __launch_bounds__(TPB) void k_aero_list(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                   int64_t n, int32_t* __restrict__ list, int32_t* counts)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    bool use = i < n && part_id[i] >= c.n_bound;
    if (use && (s.b[i] != FJSPH_FREE || s.lam_nb[i] > c.lam_cutoff)) {
        s.cellID[i] = -3;
        use = false;
    }
    const uint64_t m = __ballot(use);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int64_t g = (i - lane) / AERO_SEG; /* the wave's segment */
    int32_t at = 0;
    if (lane == leader) at = atomicAdd(counts + g * AERO_CSTRIDE, __popcll(m));
    at = __shfl(at, leader, 64);
    if (use) list[g * AERO_SEG + at + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)i;
}

/* mode 1: the points pts[n*D], outputs out[n*(D+2)]; mode 2: the slots of
 * the segmented list of k_aero_list, outputs to their cell fields. */
template <int D>
__global__ __launch_bounds__(TPB) void k_surrogate(DevState s, Surrogate sg, int64_t n, int mode,
                                                   const double* __restrict__ pts, double* __restrict__ out,
                                                   const int32_t* __restrict__ list, const int32_t* __restrict__ count)
{
    __shared__ __attribute__((aligned(16))) __bf16 lds_act[NWAVE][64 * SROW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __bf16* act = lds_act[wv];
    const int64_t base = (blockIdx.x * (int64_t)TPB) + wv * 64;
    /* mode 2: the wave's 64 list positions lie in one segment */
    const int64_t nn = mode == 2 ? (base / AERO_SEG) * AERO_SEG + count[(base / AERO_SEG) * AERO_CSTRIDE] : n;
    if (base >= nn) return;
    const bool use = base + lane < nn;
    const int64_t i = !use ? 0 : (mode == 2 ? int64_t(list[base + lane]) : base + lane);
    double x[3] = {0.0, 0.0, 0.0};
    if (use) {
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = mode == 2 ? s.x[i * D + a] : pts[i * D + a];
    }
    /* input row: normalised position, zero-padded to 32 features */
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        float v = 0.f;
        if (k < D) v = float((x[k] - sg.in_shift[k]) * sg.in_scale_inv[k]);
        act[lane * SROW + k] = __bf16(v);
    }
    __builtin_amdgcn_wave_barrier();
    f32x4 acc[4][4];
    const bf16x8* W = (const bf16x8*)sg.w;
    mlp_layer<4, 1>(W, act, acc, lane);
    store_hidden(acc, sg.b, act, lane);
    for (int l = 1; l < sg.n_hidden; ++l) {
        mlp_layer<4, 2>(W + (size_t)(4 * 1 + (l - 1) * 8) * 64, act, acc, lane);
        store_hidden(acc, sg.b + l * SW, act, lane);
    }
    mlp_layer<1, 2>(W + (size_t)(4 * 1 + (sg.n_hidden - 1) * 8) * 64, act, acc, lane);
    /* output tile: feature 4*(lane/16) + e of point 16*nt + lane%16; gather
     * the point's D+2 outputs through LDS */
    __builtin_amdgcn_wave_barrier();
    float* res = (float*)act; /* 64 points x 8 floats, reuses the rows */
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int f = 4 * (lane >> 4) + e;
            if (f < 8) res[(nt * 16 + (lane & 15)) * 8 + f] = acc[0][nt][e];
        }
    __builtin_amdgcn_wave_barrier();
    if (!use) return;
    double y[D + 2];
#pragma unroll
    for (int k = 0; k < D + 2; ++k) y[k] = double(res[lane * 8 + k] + sg.bo[k]) * sg.out_scale[k] + sg.out_shift[k];
    if (mode == 1) {
#pragma unroll
        for (int k = 0; k < D + 2; ++k) out[i * (D + 2) + k] = y[k];
        return;
    }
    s.cellID[i] = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) s.cellV[i * D + a] = y[a];
    s.cellP[i] = y[D];
    s.cellRho[i] = y[D + 1];
}

void launch_surrogate(int dim, const Consts& c, DevState s, const int32_t* part_id, Surrogate sg, int64_t n, int mode,
                      const double* pts, double* out, hipStream_t st)
{
    (void)c;
    (void)part_id;
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_surrogate<2>, g, b, 0, st, s, sg, n, mode, pts, out, nullptr, nullptr);
    else hipLaunchKernelGGL(k_surrogate<3>, g, b, 0, st, s, sg, n, mode, pts, out, nullptr, nullptr);
}

void launch_aero_list(int dim, const Consts& c, DevState s, const int32_t* part_id, int64_t n, int32_t* list,
                      int32_t* count, hipStream_t st)
{
    (void)dim;
    hipLaunchKernelGGL(k_aero_list, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, part_id, n, list, count);
}

/* The grid covers the n slots; waves past their segment's length return
 * at once. */
void launch_surrogate_list(int dim, DevState s, Surrogate sg, int64_t n, const int32_t* list, const int32_t* count,
                           hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_surrogate<2>, g, b, 0, st, s, sg, n, 2, nullptr, nullptr, list, count);
    else hipLaunchKernelGGL(k_surrogate<3>, g, b, 0, st, s, sg, n, 2, nullptr, nullptr, list, count);
}

}  // namespace fj
