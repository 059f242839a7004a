/*
 * Device helpers shared by the neighbour-build and pair-pass kernels:
 * reductions, the exact squared distance, the Wendland C2 kernel, the EOS,
 * and the Eigen-rule 2x2/3x3 linear algebra of dSPH_PreStep.
 */
#ifndef FJSPH_DEVICE_COMMON_HPP
#define FJSPH_DEVICE_COMMON_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "engine_internal.hpp"

namespace fj {

constexpr int TPB = 256;
constexpr int NWAVE = TPB / 64;
constexpr double kPi = 3.14159265358979323846;

inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

/* ------------------------------------------------------------ reductions */

struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
struct OpMin { __device__ double operator()(double a, double b) const { return a < b ? a : b; } };

/* Fixed-order block reduction; every lane receives the result. */
template <typename Op>
__device__ double block_reduce(double v, Op op, double* lds)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = lds[0];
#pragma unroll
    for (int w = 1; w < NWAVE; ++w) r = op(r, lds[w]);
    return r;
}

/* ---------------------------------------------------------------- helpers */

/* Squared distance with the reference's association and no contraction:
 * ((d0*d0 + d1*d1) + d2*d2), d = x_i - x_j (nanoflann L2_Simple). */
template <int D>
__device__ __forceinline__ double exact_dist2(const double* xi, const double* xj)
{
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double d = __dsub_rn(xi[k], xj[k]);
        acc = __dadd_rn(acc, __dmul_rn(d, d));
    }
    return acc;
}

/* Wendland C2 value and gradient factor (src/Kernel.h:37-61): GradK(R) =
 * fg(r) * R with fg = 5 Wc t^3 / H^2, t = 1 - r/(2H), zero for r/H < 1e-12. */
template <typename T>
__device__ __forceinline__ void wendland(T r, T H, T inv_H, T Wc, T& W, T& fg)
{
    const T q = r * inv_H;
    const T t = T(1) - T(0.5) * q;
    const T t2 = t * t;
    const T t3 = t2 * t;
    W = t2 * t2 * (T(2) * q + T(1)) * Wc;
    fg = (q < T(1e-12)) ? T(0) : T(5) * inv_H * inv_H * t3 * Wc;
}

template <int D>
__device__ __forceinline__ void load_vec(const double* __restrict__ a, int64_t i, double* out)
{
#pragma unroll
    for (int k = 0; k < D; ++k) out[k] = a[i * D + k];
}

template <int D>
__device__ __forceinline__ double dotd(const double* a, const double* b)
{
    double s = a[0] * b[0];
#pragma unroll
    for (int k = 1; k < D; ++k) s += a[k] * b[k];
    return s;
}

template <int D>
__device__ __forceinline__ void normalize(const double* a, double* out)
{
    const double z = dotd<D>(a, a);
    if (z > 0.0) {
        const double inv = 1.0 / sqrt(z);
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = a[k] * inv;
    } else {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = a[k];
    }
}

__device__ __forceinline__ double eos_pressure(const Consts& c, double rho)
{
    if (c.eos == FJSPH_COLE_EOS) return c.B * (pow(rho / c.rho0, c.gam) - 1.0) + c.press_back;
    return c.c0_sq * (rho - c.rho0) + c.press_back;
}

__device__ __forceinline__ double eos_density(const Consts& c, double press)
{
    if (c.eos == FJSPH_COLE_EOS) return c.rho0 * pow(((press - c.press_back) / c.B) + 1.0, 1.0 / c.gam);
    return (press - c.press_back) / c.c0_sq + c.rho0;
}

/* ColPivHouseholderQR + isInvertible + inverse with Eigen's rules
 * (pivot on the largest updated column norm, rank = #{|R_ii| >
 * maxpivot * D * eps} among the nonzero pivots). Returns false when the
 * matrix is not invertible. */
template <int D>
__device__ bool qr_inverse(double A[D][D], double inv[D][D])
{
    const double eps = 2.220446049250313e-16;
    double hc[D], nd[D], nu[D];
    int tr[D];
    double maxn = 0.0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < D; ++i) s += A[i][j] * A[i][j];
        nd[j] = sqrt(s);
        nu[j] = nd[j];
        maxn = fmax(maxn, nd[j]);
    }
    const double thr_help = (maxn * eps) * (maxn * eps) / double(D);
    const double downdate_thr = sqrt(eps);
    int nzp = D;
    double maxpivot = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        int big = k;
        double bv = nu[k];
#pragma unroll
        for (int j = k + 1; j < D; ++j)
            if (nu[j] > bv) { bv = nu[j]; big = j; }
        if (nzp == D && bv * bv < thr_help * double(D - k)) nzp = k;
        tr[k] = big;
        if (big != k) {
#pragma unroll
            for (int i = 0; i < D; ++i) { double t = A[i][k]; A[i][k] = A[i][big]; A[i][big] = t; }
            double t = nu[k]; nu[k] = nu[big]; nu[big] = t;
            t = nd[k]; nd[k] = nd[big]; nd[big] = t;
        }
        double tail = 0.0;
#pragma unroll
        for (int i = k + 1; i < D; ++i) tail += A[i][k] * A[i][k];
        const double c0 = A[k][k];
        double tau, beta;
        if (tail <= 2.2250738585072014e-308) {
            tau = 0.0;
            beta = c0;
#pragma unroll
            for (int i = k + 1; i < D; ++i) A[i][k] = 0.0;
        } else {
            beta = sqrt(c0 * c0 + tail);
            if (c0 >= 0.0) beta = -beta;
            const double den = c0 - beta;
#pragma unroll
            for (int i = k + 1; i < D; ++i) A[i][k] = A[i][k] / den;
            tau = (beta - c0) / beta;
        }
        hc[k] = tau;
        A[k][k] = beta;
        maxpivot = fmax(maxpivot, fabs(beta));
        if (tau != 0.0) {
#pragma unroll
            for (int j = k + 1; j < D; ++j) {
                double tmp = 0.0;
#pragma unroll
                for (int i = k + 1; i < D; ++i) tmp += A[i][k] * A[i][j];
                tmp += A[k][j];
                A[k][j] -= tau * tmp;
#pragma unroll
                for (int i = k + 1; i < D; ++i) A[i][j] -= tau * A[i][k] * tmp;
            }
        }
#pragma unroll
        for (int j = k + 1; j < D; ++j) {
            if (nu[j] != 0.0) {
                double t = fabs(A[k][j]) / nu[j];
                t = (1.0 + t) * (1.0 - t);
                t = t < 0.0 ? 0.0 : t;
                const double ratio = nu[j] / nd[j];
                if (t * ratio * ratio <= downdate_thr) {
                    double s = 0.0;
#pragma unroll
                    for (int i = k + 1; i < D; ++i) s += A[i][j] * A[i][j];
                    nd[j] = sqrt(s);
                    nu[j] = nd[j];
                } else {
                    nu[j] *= sqrt(t);
                }
            }
        }
    }
    const double thr = maxpivot * (double(D) * eps);
    int rank = 0;
#pragma unroll
    for (int i = 0; i < D; ++i)
        if (i < nzp && fabs(A[i][i]) > thr) ++rank;
    if (rank != D) return false;
    double cm[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) cm[i][j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double tau = hc[k];
        if (tau == 0.0) continue;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double tmp = cm[k][j];
#pragma unroll
            for (int i = k + 1; i < D; ++i) tmp += A[i][k] * cm[i][j];
            cm[k][j] -= tau * tmp;
#pragma unroll
            for (int i = k + 1; i < D; ++i) cm[i][j] -= tau * A[i][k] * tmp;
        }
    }
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
        for (int i = D - 1; i >= 0; --i) {
            double s = cm[i][j];
#pragma unroll
            for (int k = i + 1; k < D; ++k) s -= A[i][k] * cm[k][j];
            cm[i][j] = s / A[i][i];
        }
    int perm[D];
#pragma unroll
    for (int i = 0; i < D; ++i) perm[i] = i;
#pragma unroll
    for (int k = 0; k < D; ++k) { int t = perm[k]; perm[k] = perm[tr[k]]; perm[tr[k]] = t; }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) inv[perm[i]][j] = cm[i][j];
    return true;
}

/* SelfAdjointEigenSolver::computeDirect smallest eigenvalue (lower triangle). */
inline __device__ double min_eig(const double m[2][2])
{
    const double shift = (m[0][0] + m[1][1]) / 2.0;
    double s00 = m[0][0] - shift, s11 = m[1][1] - shift, s10 = m[1][0];
    const double scale = fmax(fabs(s00), fmax(fabs(s11), fabs(s10)));
    if (scale > 0.0) { s00 /= scale; s11 /= scale; s10 /= scale; }
    const double t0 = 0.5 * sqrt((s00 - s11) * (s00 - s11) + 4.0 * (s10 * s10));
    const double t1 = 0.5 * (s00 + s11);
    return (t1 - t0) * scale + shift;
}

inline __device__ double min_eig(const double a[3][3])
{
    const double shift = (a[0][0] + a[1][1] + a[2][2]) / 3.0;
    double m[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) m[i][j] = (i >= j) ? a[i][j] : a[j][i];
#pragma unroll
    for (int i = 0; i < 3; ++i) m[i][i] -= shift;
    double scale = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) scale = fmax(scale, fabs(m[i][j]));
    if (scale > 0.0) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) m[i][j] /= scale;
    }
    const double inv3 = 1.0 / 3.0;
    const double c0 = m[0][0] * m[1][1] * m[2][2] + 2.0 * m[1][0] * m[2][0] * m[2][1] - m[0][0] * m[2][1] * m[2][1] -
                      m[1][1] * m[2][0] * m[2][0] - m[2][2] * m[1][0] * m[1][0];
    const double c1 = m[0][0] * m[1][1] - m[1][0] * m[1][0] + m[0][0] * m[2][2] - m[2][0] * m[2][0] +
                      m[1][1] * m[2][2] - m[2][1] * m[2][1];
    const double c2 = m[0][0] + m[1][1] + m[2][2];
    const double c2o3 = c2 * inv3;
    double ao3 = (c2 * c2o3 - c1) * inv3;
    ao3 = fmax(ao3, 0.0);
    const double hb = 0.5 * (c0 + c2o3 * (2.0 * c2o3 * c2o3 - c1));
    double q = ao3 * ao3 * ao3 - hb * hb;
    q = fmax(q, 0.0);
    const double rho = sqrt(ao3);
    const double th = atan2(sqrt(q), hb) * inv3;
    const double ct = cos(th), st = sin(th);
    const double s3 = sqrt(3.0);
    const double r0 = c2o3 - rho * (ct + s3 * st);
    const double r1 = c2o3 - rho * (ct - s3 * st);
    const double r2 = c2o3 + 2.0 * rho * ct;
    return fmin(r0 * scale + shift, fmin(r1 * scale + shift, r2 * scale + shift));
}


template <typename T>
__device__ __forceinline__ T fast_rcp(T x) { return T(1) / x; }
template <>
__device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

/* Square root and cosine of the FP32 pair terms: the hardware instructions
 * (v_sqrt_f32, about 1 ulp; v_cos_f32, absolute error near 1e-6 for the
 * arguments in [0, 3pi/2] used here) instead of the correctly rounded
 * library sequences.  FP64: the library functions. */
template <typename T>
__device__ __forceinline__ T fast_sqrt(T x) { return sqrt(x); }
template <>
__device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename T>
__device__ __forceinline__ T fast_cos(T x) { return cos(x); }
template <>
__device__ __forceinline__ float fast_cos<float>(float x) { return __cosf(x); }

/* Determinant of the 4x4 matrix with rows (p_k, 1): Laplace expansion in
 * 2x2 minors of columns 0-1 (the signed volume of the mesh crossing tests;
 * contraction off, so its sign agrees with the CPU restatement). */
__device__ __forceinline__ double det4_rows(const double* p0, const double* p1, const double* p2, const double* p3)
{
#pragma clang fp contract(off)
    const double* r[4] = {p0, p1, p2, p3};
    double m[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        m[i][0] = r[i][0];
        m[i][1] = r[i][1];
        m[i][2] = r[i][2];
        m[i][3] = 1.0;
    }
    auto d2 = [&](int i0, int i1) { return m[i0][0] * m[i1][1] - m[i1][0] * m[i0][1]; };
    const double d2_01 = d2(0, 1), d2_02 = d2(0, 2), d2_03 = d2(0, 3);
    const double d2_12 = d2(1, 2), d2_13 = d2(1, 3), d2_23 = d2(2, 3);
    auto d3 = [&](int i0, double a0, int i1, double a1, int i2, double a2) {
        return m[i0][2] * a0 + (-m[i1][2] * a1 + m[i2][2] * a2);
    };
    const double d3_0 = d3(1, d2_23, 2, d2_13, 3, d2_12);
    const double d3_1 = d3(0, d2_23, 2, d2_03, 3, d2_02);
    const double d3_2 = d3(0, d2_13, 1, d2_03, 3, d2_01);
    const double d3_3 = d3(0, d2_12, 1, d2_02, 2, d2_01);
    return (-m[0][3] * d3_0 + m[1][3] * d3_1) + (-m[2][3] * d3_2 + m[3][3] * d3_3);
}

}  // namespace fj

#endif
