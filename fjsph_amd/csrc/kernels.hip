/*
 * gfx950 kernels of the WCSPH step engine.
 *
 * One lane per particle slot; slots are sorted by cell so a wavefront's
 * neighbours sit in a few contiguous strips and the neighbour gathers hit L2.
 * Pair terms never scatter: every pass is owner-computes over the particle's
 * own list, so results do not depend on launch order.  Reductions are
 * deterministic (fixed-order wave / block / two-stage trees).
 *
 * Reference correspondence (file:line):
 *   k_cells/k_scan/k_scatter/k_sort_cells/k_build_lists  update_neighbours, src/Neighbours.cpp:7-47
 *   k_dsph         dSPH_PreStep              src/Shifting.cpp:12-123
 *   k_aero_vel     get_aero_velocity const   src/Resid.cpp:569-611
 *   k_surface1     Detect_Surface loop 1     src/Geometry.cpp:26-143
 *                  + dissipation_terms       src/Shifting.cpp:126-186
 *   k_surface2     Detect_Surface loop 2/3   src/Geometry.cpp:145-277
 *   k_shift        particle_shift            src/Shifting.cpp:189-290
 *   k_bnd_*        boundary solvers          src/Resid.cpp:21-186, src/Newmark_Beta.cpp:69-132
 *   k_force        get_acc_and_Rrho_on_i     src/Resid.cpp:243-418
 *                  + Do_NB_Iter update       src/Newmark_Beta.cpp:137-240
 *   k_dt_reduce    find_timestep reduction   src/Integration.cpp:384-403
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "engine_internal.hpp"

namespace fj {

constexpr int TPB = 256;
constexpr int NWAVE = TPB / 64;
constexpr double kPi = 3.14159265358979323846;

static inline unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }
int reduction_parts(int64_t n) { return (int)grid_of(n); }

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
__device__ double min_eig(const double m[2][2])
{
    const double shift = (m[0][0] + m[1][1]) / 2.0;
    double s00 = m[0][0] - shift, s11 = m[1][1] - shift, s10 = m[1][0];
    const double scale = fmax(fabs(s00), fmax(fabs(s11), fabs(s10)));
    if (scale > 0.0) { s00 /= scale; s11 /= scale; s10 /= scale; }
    const double t0 = 0.5 * sqrt((s00 - s11) * (s00 - s11) + 4.0 * (s10 * s10));
    const double t1 = 0.5 * (s00 + s11);
    return (t1 - t0) * scale + shift;
}

__device__ double min_eig(const double a[3][3])
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

/* ------------------------------------------------------- neighbour build */

template <int D>
__global__ __launch_bounds__(TPB) void k_bbox(const double* __restrict__ x, int64_t n, double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double v = i < n ? x[i * D + k] : x[k];
        const double lo = block_reduce(v, OpMin(), lds);
        const double hi = block_reduce(v, OpMax(), lds);
        if (threadIdx.x == 0) {
            partial[blockIdx.x * 2 * D + k] = lo;
            partial[blockIdx.x * 2 * D + D + k] = hi;
        }
    }
}

template <int D>
__global__ __launch_bounds__(TPB) void k_bbox_final(const double* partial, int nparts, double* out)
{
    __shared__ double lds[NWAVE];
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) {
        const bool is_min = k < D;
        double v = partial[k];
        for (int b = threadIdx.x; b < nparts; b += TPB) {
            const double w = partial[b * 2 * D + k];
            v = is_min ? fmin(v, w) : fmax(v, w);
        }
        const double r = is_min ? block_reduce(v, OpMin(), lds) : block_reduce(v, OpMax(), lds);
        if (threadIdx.x == 0) out[k] = r;
    }
}

void launch_bbox(int dim, const double* x, int64_t n, double* partial, int nparts, double* out, hipStream_t st)
{
    if (dim == 2) {
        hipLaunchKernelGGL(k_bbox<2>, dim3(nparts), dim3(TPB), 0, st, x, n, partial);
        hipLaunchKernelGGL(k_bbox_final<2>, dim3(1), dim3(TPB), 0, st, partial, nparts, out);
    } else {
        hipLaunchKernelGGL(k_bbox<3>, dim3(nparts), dim3(TPB), 0, st, x, n, partial);
        hipLaunchKernelGGL(k_bbox_final<3>, dim3(1), dim3(TPB), 0, st, partial, nparts, out);
    }
}

struct Grid3 {
    double lo[3];
    double inv_cs;
    int32_t dims[3];
};

template <int D>
__global__ __launch_bounds__(TPB) void k_cells(const double* __restrict__ x, int64_t n, Grid3 g, int32_t* cell,
                                               int32_t* cell_count)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    int32_t ci[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double f = floor((x[i * D + k] - g.lo[k]) * g.inv_cs);
        int32_t c = (int32_t)fmin(fmax(f, 0.0), double(g.dims[k] - 1));
        ci[k] = c;
    }
    const int32_t lin = (D == 3) ? ((ci[2] * g.dims[1] + ci[1]) * g.dims[0] + ci[0]) : (ci[1] * g.dims[0] + ci[0]);
    cell[i] = lin;
    atomicAdd(&cell_count[lin], 1);
}

void launch_cells(int dim, const double* x, int64_t n, const double* lo, double inv_cs, const int32_t* dims,
                  int32_t* cell, int32_t* cell_count, hipStream_t st)
{
    Grid3 g;
    for (int k = 0; k < 3; ++k) { g.lo[k] = k < dim ? lo[k] : 0.0; g.dims[k] = dims[k]; }
    g.inv_cs = inv_cs;
    if (dim == 2) hipLaunchKernelGGL(k_cells<2>, dim3(grid_of(n)), dim3(TPB), 0, st, x, n, g, cell, cell_count);
    else hipLaunchKernelGGL(k_cells<3>, dim3(grid_of(n)), dim3(TPB), 0, st, x, n, g, cell, cell_count);
}

/* Three-phase exclusive scan over tiles of SCAN_TILE counts. */
constexpr int SCAN_TILE = 2048;

__device__ int32_t block_excl_scan_i(int32_t v, int32_t* lds, int32_t& total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    __syncthreads();
    if (lane == 63) lds[wid] = incl;
    __syncthreads();
    int32_t woff = 0;
    total = 0;
    for (int w = 0; w < NWAVE; ++w) {
        if (w < wid) woff += lds[w];
        total += lds[w];
    }
    return woff + incl - v;
}

__global__ __launch_bounds__(TPB) void k_scan_tiles(const int32_t* __restrict__ count, int64_t ncell, int32_t* tile_sums)
{
    __shared__ int32_t lds[NWAVE];
    const int64_t base = blockIdx.x * (int64_t)SCAN_TILE;
    int32_t s = 0;
    for (int k = threadIdx.x; k < SCAN_TILE; k += TPB) {
        const int64_t c = base + k;
        if (c < ncell) s += count[c];
    }
    int32_t total;
    (void)block_excl_scan_i(s, lds, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(TPB) void k_scan_sums(int32_t* tile_sums, int ntiles)
{
    __shared__ int32_t lds[NWAVE];
    int32_t carry = 0;
    for (int base = 0; base < ntiles; base += TPB) {
        const int t = base + threadIdx.x;
        const int32_t v = t < ntiles ? tile_sums[t] : 0;
        int32_t total;
        const int32_t ex = block_excl_scan_i(v, lds, total);
        if (t < ntiles) tile_sums[t] = carry + ex;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(TPB) void k_scan_apply(const int32_t* __restrict__ count, int64_t ncell,
                                                    const int32_t* __restrict__ tile_sums, int32_t* start)
{
    __shared__ int32_t lds[NWAVE];
    constexpr int PER = SCAN_TILE / TPB;
    const int64_t base = blockIdx.x * (int64_t)SCAN_TILE + threadIdx.x * PER;
    int32_t v[PER];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = (base + k < ncell) ? count[base + k] : 0;
        s += v[k];
    }
    int32_t total;
    int32_t ex = block_excl_scan_i(s, lds, total) + tile_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (base + k < ncell) start[base + k] = ex;
        ex += v[k];
    }
}

__global__ void k_scan_total(const int32_t* count, int64_t ncell, int32_t* start)
{
    start[ncell] = start[ncell - 1] + count[ncell - 1];
}

void launch_scan(const int32_t* count, int64_t ncell, int32_t* start, int32_t* tile_sums, hipStream_t st)
{
    const int ntiles = (int)((ncell + SCAN_TILE - 1) / SCAN_TILE);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(TPB), 0, st, count, ncell, tile_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(TPB), 0, st, tile_sums, ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(TPB), 0, st, count, ncell, tile_sums, start);
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, st, count, ncell, start);
}

__global__ __launch_bounds__(TPB) void k_scatter(const int32_t* __restrict__ cell, int64_t n,
                                                 const int32_t* __restrict__ start, int32_t* fill, int32_t* order)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cell[i];
    const int32_t pos = start[c] + atomicAdd(&fill[c], 1);
    order[pos] = (int32_t)i;
}

void launch_scatter(const int32_t* cell, int64_t n, const int32_t* start, int32_t* fill, int32_t* order,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(TPB), 0, st, cell, n, start, fill, order);
}

/* Sort each cell's slot list ascending so list order is deterministic: one
 * workgroup per cell, segment staged in LDS, each lane writes its elements
 * at their rank.  Segments longer than the LDS tile fall back to an
 * insertion sort in one lane. */
constexpr int SORT_TILE = 1024;

__global__ __launch_bounds__(64) void k_sort_cells(const int32_t* __restrict__ start, int64_t ncell, int32_t* order)
{
    __shared__ int32_t seg[SORT_TILE];
    const int64_t c = blockIdx.x;
    if (c >= ncell) return;
    const int32_t a = start[c], b = start[c + 1];
    const int32_t k = b - a;
    if (k <= 1) return;
    if (k > SORT_TILE) {
        if (threadIdx.x == 0) {
            for (int32_t i = a + 1; i < b; ++i) {
                const int32_t v = order[i];
                int32_t j = i - 1;
                while (j >= a && order[j] > v) { order[j + 1] = order[j]; --j; }
                order[j + 1] = v;
            }
        }
        return;
    }
    for (int t = threadIdx.x; t < k; t += 64) seg[t] = order[a + t];
    __syncthreads();
    for (int t = threadIdx.x; t < k; t += 64) {
        const int32_t v = seg[t];
        int rank = 0;
        for (int u = 0; u < k; ++u) rank += seg[u] < v;
        order[a + rank] = v;
    }
}

void launch_sort_cells(const int32_t* start, int64_t ncell, int32_t* order, hipStream_t st)
{
    hipLaunchKernelGGL(k_sort_cells, dim3((unsigned)ncell), dim3(64), 0, st, start, ncell, order);
}

/* Positions in cell order (xs[s] = x[order[s]]) so the list build streams
 * candidate strips contiguously. */
template <int D>
__global__ __launch_bounds__(TPB) void k_gather_x(const double* __restrict__ x, const int32_t* __restrict__ order,
                                                  int64_t n, double* xs)
{
    const int64_t s = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (s >= n) return;
    const int64_t j = order[s];
#pragma unroll
    for (int k = 0; k < D; ++k) xs[s * D + k] = x[j * D + k];
}

void launch_gather_x(int dim, const double* x, const int32_t* order, int64_t n, double* xs, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_gather_x<2>, dim3(grid_of(n)), dim3(TPB), 0, st, x, order, n, xs);
    else hipLaunchKernelGGL(k_gather_x<3>, dim3(grid_of(n)), dim3(TPB), 0, st, x, order, n, xs);
}

/* Neighbour lists: for each slot, the 3^(D-1) x-strips of the stencil
 * (3 x-adjacent cells are contiguous in cell order), exact FP64 predicate
 * d2 < sr, self excluded.  Counts beyond the capacity are still counted so
 * the host can grow the lists and rebuild. */
template <int D, typename RT>
__global__ __launch_bounds__(TPB) void k_build_lists(const double* __restrict__ x, const double* __restrict__ xs,
                                                     int64_t n, const int32_t* __restrict__ cell, Grid3 g,
                                                     const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ order, double sr, NbrList nl,
                                                     DevScalars* sc)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    double xi[D];
    load_vec<D>(x, i, xi);
    const int32_t lin = cell[i];
    const int32_t cx = lin % g.dims[0];
    const int32_t cy = (lin / g.dims[0]) % g.dims[1];
    const int32_t cz = (D == 3) ? lin / (g.dims[0] * g.dims[1]) : 0;
    const int32_t x0 = max(cx - 1, 0), x1 = min(cx + 1, g.dims[0] - 1);
    RT* pay = (RT*)nl.r;
    int cnt = 0;
    const int zlo = (D == 3) ? max(cz - 1, 0) : 0, zhi = (D == 3) ? min(cz + 1, g.dims[2] - 1) : 0;
    for (int z = zlo; z <= zhi; ++z) {
        for (int y = max(cy - 1, 0); y <= min(cy + 1, g.dims[1] - 1); ++y) {
            const int32_t row = (D == 3) ? (z * g.dims[1] + y) * g.dims[0] : y * g.dims[0];
            const int32_t a = start[row + x0], b = start[row + x1 + 1];
            for (int32_t s = a; s < b; ++s) {
                double xj[D];
                load_vec<D>(xs, s, xj);
                const double d2 = exact_dist2<D>(xi, xj);
                if (d2 < sr) {
                    const int32_t j = order[s];
                    if (j == (int32_t)i) continue;
                    if (cnt < nl.cap) {
                        nl.idx[(int64_t)cnt * nl.npad + i] = j;
                        pay[(int64_t)cnt * nl.npad + i] = (sizeof(RT) == 4) ? RT(sqrt(d2)) : RT(d2);
                    }
                    ++cnt;
                }
            }
        }
    }
    nl.count[i] = cnt;
    atomicMax(&sc->max_count, cnt);
}

void launch_build_lists(int dim, const double* x, const double* xs, int64_t n, const int32_t* cell,
                        const int32_t* dims, const int32_t* start, const int32_t* order, double sr, NbrList nl,
                        DevScalars* sc, hipStream_t st)
{
    Grid3 g;
    for (int k = 0; k < 3; ++k) { g.lo[k] = 0.0; g.dims[k] = dims[k]; }
    g.inv_cs = 0.0;
    const dim3 gr(grid_of(n)), bl(TPB);
    if (dim == 2) {
        if (nl.f32) hipLaunchKernelGGL((k_build_lists<2, float>), gr, bl, 0, st, x, xs, n, cell, g, start, order, sr, nl, sc);
        else hipLaunchKernelGGL((k_build_lists<2, double>), gr, bl, 0, st, x, xs, n, cell, g, start, order, sr, nl, sc);
    } else {
        if (nl.f32) hipLaunchKernelGGL((k_build_lists<3, float>), gr, bl, 0, st, x, xs, n, cell, g, start, order, sr, nl, sc);
        else hipLaunchKernelGGL((k_build_lists<3, double>), gr, bl, 0, st, x, xs, n, cell, g, start, order, sr, nl, sc);
    }
}

/* Per-particle derived quantities (see Aux). */
template <int D>
__global__ __launch_bounds__(TPB) void k_aux(DevState s, Aux a, int64_t n, int ale, int mixed)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    const double rho = s.rho[i], m = s.m[i], p = s.p[i];
    const double vol = m / rho, P = p / (rho * rho), ir = 1.0 / rho;
    a.vol[i] = vol;
    a.P[i] = P;
    a.invrho[i] = ir;
    if (mixed) {
        const float ms = s.b[i] == FJSPH_BOUND ? -float(m) : float(m);
        a.fq[i] = make_float4(ms, float(P), float(ir), float(vol));
        if (ale) {
            float vp[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < D; ++k) vp[k] = float(s.vPert[i * D + k]);
            a.fale[i] = make_float4(vp[0], vp[1], vp[2], float(rho));
        }
    }
}

void launch_aux(int dim, DevState s, Aux a, int64_t n, int ale, int mixed, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_aux<2>, dim3(grid_of(n)), dim3(TPB), 0, st, s, a, n, ale, mixed);
    else hipLaunchKernelGGL(k_aux<3>, dim3(grid_of(n)), dim3(TPB), 0, st, s, a, n, ale, mixed);
}

__global__ __launch_bounds__(TPB) void k_gather_d(const double* __restrict__ src, double* __restrict__ dst,
                                                  const int32_t* __restrict__ perm, int64_t n, int width)
{
    const int64_t t = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (t >= n * width) return;
    const int64_t s = t / width;
    const int k = (int)(t - s * width);
    dst[t] = src[(int64_t)perm[s] * width + k];
}

__global__ __launch_bounds__(TPB) void k_gather_i(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                                  const int32_t* __restrict__ perm, int64_t n)
{
    const int64_t s = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (s >= n) return;
    dst[s] = src[perm[s]];
}

void launch_gather_d(const double* src, double* dst, const int32_t* perm, int64_t n, int width, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_d, dim3(grid_of(n * width)), dim3(TPB), 0, st, src, dst, perm, n, width);
}

__global__ __launch_bounds__(TPB) void k_iota(int32_t* a, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

void launch_iota(int32_t* a, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(TPB), 0, st, a, n);
}

void launch_gather_i(const int32_t* src, int32_t* dst, const int32_t* perm, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_i, dim3(grid_of(n)), dim3(TPB), 0, st, src, dst, perm, n);
}

/* ------------------------------------------------------------ pair helpers */

/* The distance of pair k of slot i as the reference uses it: from the d^2
 * frozen at the last neighbour build (src/Resid.cpp:288-289).  When the
 * positions are still those of the build (r_from_x) it is recomputed from x
 * with the build's own arithmetic, which gives the same bits. */
template <int D>
__device__ __forceinline__ void pair_dist(int r_from_x, const double* xi, const double* xj, const NbrList& nl,
                                          int k, int64_t i, double& d2, double& r)
{
    if (r_from_x) {
        d2 = exact_dist2<D>(xi, xj);
        r = sqrt(d2);
    } else if (nl.f32) {
        r = double(((const float*)nl.r)[(int64_t)k * nl.npad + i]);
        d2 = r * r;
    } else {
        d2 = ((const double*)nl.r)[(int64_t)k * nl.npad + i];
        r = sqrt(d2);
    }
}

template <typename T>
__device__ __forceinline__ T fast_rcp(T x) { return T(1) / x; }
template <>
__device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

/* ------------------------------------------------------------ dSPH_PreStep */

template <int D>
__device__ __forceinline__ void aero_velocity_one(const Consts& c, DevState& s, int64_t i, double lam_nb, int cnt_with_self)
{
    const int32_t b = s.b[i];
    const bool cond = c.use_lam ? (lam_nb < c.lam_cutoff && b == FJSPH_FREE)
                                : (double(cnt_with_self) * c.i_n_full < c.lam_cutoff && b == FJSPH_FREE);
#pragma unroll
    for (int k = 0; k < D; ++k) s.cellV[i * D + k] = cond ? c.v_inf[k] : 0.0;
    s.cellID[i] = cond ? 1 : -3;
}

template <int D>
__global__ __launch_bounds__(TPB) void k_dsph(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                              NbrList nl, int64_t n, int r_from_x, int with_aero, double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    double npd_i = 0.0;
    if (i < n) {
        double xi[D];
        load_vec<D>(s.x, i, xi);
        const double rhoi = s.rho[i];
        double Lm[D][D], Lnb[D][D], gR[D], nr[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            gR[a] = 0.0;
            nr[a] = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) { Lm[a][b2] = 0.0; Lnb[a][b2] = 0.0; }
        }
        double kernsum = c.Wc, colour = 0.0;
        const double inv_H = 1.0 / c.H;
        const int cnt = nl.count[i];
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
            double xj[D], d2, r;
            load_vec<D>(s.x, j, xj);
            pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
            double Rji[D];
#pragma unroll
            for (int a = 0; a < D; ++a) Rji[a] = xj[a] - xi[a];
            const double volj = ax.vol[j];
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            double G[D];
#pragma unroll
            for (int a = 0; a < D; ++a) G[a] = -fg * Rji[a];
            const double drho = volj * (s.rho[j] - rhoi);
            const bool nb = s.b[j] > FJSPH_PISTON;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const double vr = volj * Rji[a];
#pragma unroll
                for (int b2 = 0; b2 < D; ++b2) {
                    const double t = vr * G[b2];
                    Lm[a][b2] -= t;
                    if (nb) Lnb[a][b2] -= t;
                }
                gR[a] += drho * G[a];
            }
            if (nb) {
#pragma unroll
                for (int a = 0; a < D; ++a) nr[a] += volj * G[a];
                kernsum += W;
                colour += volj * W;
                npd_i += W;
            }
        }
        double Lw[D][D], Linv[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) { Lw[a][b2] = Lm[a][b2]; Linv[a][b2] = (a == b2) ? 1.0 : 0.0; }
        double tmp[D][D];
        if (qr_inverse<D>(Lw, tmp)) {
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b2 = 0; b2 < D; ++b2) Linv[a][b2] = tmp[a][b2];
        }
        double lam, lam_nb;
        if (s.b[i] == FJSPH_BOUND) {
            lam = 1.0;
            lam_nb = 1.0;
        } else {
            lam = min_eig(Lm);
            lam_nb = min_eig(Lnb);
        }
        s.lam[i] = lam;
        s.lam_nb[i] = lam_nb;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double g = 0.0, q = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) {
                g += Linv[a][b2] * gR[b2];
                q += Linv[a][b2] * nr[b2];
                s.L[i * D * D + a * D + b2] = Linv[a][b2];
            }
            s.gradRho[i * D + a] = g;
            s.norm[i * D + a] = q;
        }
        s.colourG[i] = lam > 0.7 ? 2.0 : 2.0 * fmax(1.0, 1.0 / (2.0 * colour));
        s.kernsum[i] = kernsum;
        s.colour[i] = colour;
        if (with_aero && c.aero_source == FJSPH_CONST_VEL && part_id[i] >= c.n_bound)
            aero_velocity_one<D>(c, s, i, lam_nb, cnt + 1);
    }
    const double bs = block_reduce(npd_i, OpSum(), lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

/* Sum `width` columns of per-block partials in a fixed order. */
__global__ __launch_bounds__(TPB) void k_sum_partials(const double* partial, int nparts, int width, double scale,
                                                      double* out)
{
    __shared__ double lds[NWAVE];
    for (int w = 0; w < width; ++w) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nparts; b += TPB) v += partial[(int64_t)b * width + w];
        const double r = block_reduce(v, OpSum(), lds);
        if (threadIdx.x == 0) out[w] = r * scale;
    }
}

void launch_sum_partials(const double* partial, int nparts, int width, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(TPB), 0, st, partial, nparts, width, 1.0, out);
}

void launch_dsph(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, NbrList nl, int64_t n,
                 int r_from_x, int with_aero_vel, double* partial, int nparts, DevScalars* sc, hipStream_t st)
{
    if (dim == 2)
        hipLaunchKernelGGL(k_dsph<2>, dim3(nparts), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x, with_aero_vel, partial);
    else
        hipLaunchKernelGGL(k_dsph<3>, dim3(nparts), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x, with_aero_vel, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(TPB), 0, st, partial, nparts, 1, 1.0 / double(n), &sc->npd);
}

template <int D>
__global__ __launch_bounds__(TPB) void k_aero_vel(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                  NbrList nl, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    aero_velocity_one<D>(c, s, i, s.lam_nb[i], nl.count[i] + 1);
}

void launch_aero_velocity(int dim, const Consts& c, DevState s, const int32_t* part_id, NbrList nl, int64_t n,
                          hipStream_t st)
{
    if (c.aero_source != FJSPH_CONST_VEL) return;
    if (dim == 2) hipLaunchKernelGGL(k_aero_vel<2>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, part_id, nl, n);
    else hipLaunchKernelGGL(k_aero_vel<3>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, part_id, nl, n);
}

/* -------------------------------------- Detect_Surface + dissipation_terms */

/* One dissipation_terms pair term (src/Shifting.cpp:141-173) in T. */
template <int D, typename T>
struct DissAcc {
    double av[D];
    double rd;
    T Bg, cfac, H, eps2, alpha;
    __device__ void init(const Consts& c, double rhoi)
    {
#pragma unroll
        for (int a = 0; a < D; ++a) av[a] = 0.0;
        rd = 0.0;
        Bg = T(c.B * c.gam);
        cfac = T(sqrt(c.B * c.gam / rhoi));
        H = T(c.H);
        eps2 = T(0.0001 * c.H_sq);
        alpha = T(c.visc_alpha);
    }
    __device__ void add(const DevState& s, const Aux& ax, int32_t j, const double* xi, const double* xj,
                        const double* vi, const double* gri, double rhoi, double d2, T r, T fg)
    {
        T Rji[D], Vji[D], g[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            Rji[a] = T(xj[a] - xi[a]);
            Vji[a] = T(s.v[(int64_t)j * D + a] - vi[a]);
            g[a] = fg * Rji[a];
        }
        (void)r;
        const T idist2 = fast_rcp<T>(T(d2) + eps2);
        const double rhoj = s.rho[j];
        const T volj = T(ax.vol[j]);
        T rdotg = Rji[0] * g[0];
#pragma unroll
        for (int a = 1; a < D; ++a) rdotg += Rji[a] * g[a];
        const T drho = T(rhoj - rhoi);
        if (s.b[j] > FJSPH_PISTON) {
            T vdotr = Vji[0] * Rji[0];
#pragma unroll
            for (int a = 1; a < D; ++a) vdotr += Vji[a] * Rji[a];
            if (!(vdotr > T(0))) {
                const T muij = H * vdotr * idist2;
                const T cbar = T(0.5) * (cfac + sqrt(Bg * T(ax.invrho[j])));
                const T f = T(s.m[j]) * alpha * cbar * muij * fast_rcp<T>(T(0.5 * (rhoi + rhoj)));
#pragma unroll
                for (int a = 0; a < D; ++a) av[a] += double(g[a] * f);
            }
            T gsum = T(0);
#pragma unroll
            for (int a = 0; a < D; ++a) gsum += T(gri[a] + s.gradRho[(int64_t)j * D + a]) * Rji[a];
            rd += double(volj * (drho + T(0.5) * gsum) * rdotg * idist2);
        } else {
            rd += double(volj * drho * rdotg * idist2);
        }
    }
    __device__ void store(const Consts& c, DevState& s, int64_t i)
    {
#pragma unroll
        for (int a = 0; a < D; ++a) s.aVisc[i * D + a] = av[a];
        s.deltaD[i] = c.dsph_cont * rd;
    }
};

template <int D, typename T>
__global__ __launch_bounds__(TPB) void k_surface1(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                  NbrList nl, int64_t n, int r_from_x, double* norms_tmp,
                                                  int with_diss)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    if (part_id[i] < c.n_bound) {
#pragma unroll
        for (int a = 0; a < D; ++a) norms_tmp[i * D + a] = 0.0;
        return;
    }
    const double h = 1.33 * c.particle_step;
    double xi[D], ni[D], nhat[D];
    load_vec<D>(s.x, i, xi);
    load_vec<D>(s.norm, i, ni);
    normalize<D>(ni, nhat);
    const int32_t bi = s.b[i];
    const double lam_nb = s.lam_nb[i];
    const double lami = s.lam[i];
    const int cnt = nl.count[i];
    const double inv_H = 1.0 / c.H;
    if (bi < FJSPH_PIPE) {
        s.surf[i] = 0;
        s.woccl[i] = 1.0;
    } else if (lam_nb < 0.2) {
        s.surf[i] = 1;
    } else if (lam_nb < 0.75) {
        /* Marrone et al. (2010) point-T test, src/Geometry.cpp:43-98 */
        double pT[D];
#pragma unroll
        for (int a = 0; a < D; ++a) pT[a] = xi[a] + h * nhat[a];
        int surf = 1;
        const double r_lim = sqrt(2.0) * h;
        double tau_hat[D];
        if (D == 2) {
            double tau[D];
            tau[0] = ni[1 % D];
            tau[1 % D] = -ni[0];
            normalize<D>(tau, tau_hat);
        }
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
            double xj[D], xjT[D], d2, r;
            load_vec<D>(s.x, j, xj);
            pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
#pragma unroll
            for (int a = 0; a < D; ++a) xjT[a] = xj[a] - pT[a];
            if (r >= r_lim) {
                if (sqrt(dotd<D>(xjT, xjT)) < h) { surf = 0; break; }
            } else {
                if (D == 2) {
                    if (fabs(dotd<D>(nhat, xjT)) + fabs(dotd<D>(tau_hat, xjT)) < h) { surf = 0; break; }
                } else {
                    double mr[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) mr[a] = -(xi[a] - xj[a]) / r;
                    if (acos(dotd<D>(nhat, mr)) < kPi / 4.0) { surf = 0; break; }
                }
            }
        }
        s.surf[i] = surf;
    } else {
        s.surf[i] = 0;
    }
    /* normal from the eigenvalue gradient (src/Geometry.cpp:109-141), with
     * the dissipation terms in the same neighbour loop */
    double nrm[D];
#pragma unroll
    for (int a = 0; a < D; ++a) nrm[a] = 0.0;
    const bool high = lami > 0.7;
    double vi[D], gri[D];
    DissAcc<D, T> dacc;
    const double rhoi = s.rho[i];
    if (with_diss) {
        load_vec<D>(s.v, i, vi);
        load_vec<D>(s.gradRho, i, gri);
        dacc.init(c, rhoi);
    }
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
        double xj[D], d2, r;
        load_vec<D>(s.x, j, xj);
        pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
        double W, fg;
        wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
        const double f = ax.vol[j] * (high ? (s.lam[j] - lami) : s.lam[j]) * fg;
#pragma unroll
        for (int a = 0; a < D; ++a) nrm[a] += f * (xi[a] - xj[a]);
        if (with_diss) {
            T Wt, fgt;
            if (sizeof(T) == 8) { Wt = T(W); fgt = T(fg); }
            else wendland<T>(T(r), T(c.H), T(inv_H), T(c.Wc), Wt, fgt);
            dacc.add(s, ax, j, xi, xj, vi, gri, rhoi, d2, T(r), fgt);
        }
    }
    double Ln[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b2 = 0; b2 < D; ++b2) t += s.L[i * D * D + a * D + b2] * nrm[b2];
        Ln[a] = t;
    }
    double out[D];
    if (sqrt(dotd<D>(Ln, Ln)) > 0.1 * lami / c.H) normalize<D>(Ln, out);
    else {
#pragma unroll
        for (int a = 0; a < D; ++a) out[a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) norms_tmp[i * D + a] = out[a];
    if (with_diss) dacc.store(c, s, i);
}

template <int D, typename T>
__global__ __launch_bounds__(TPB) void k_dissipation(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                     NbrList nl, int64_t n, int r_from_x)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    double xi[D], vi[D], gri[D];
    load_vec<D>(s.x, i, xi);
    load_vec<D>(s.v, i, vi);
    load_vec<D>(s.gradRho, i, gri);
    const double rhoi = s.rho[i];
    DissAcc<D, T> dacc;
    dacc.init(c, rhoi);
    const T inv_H = T(1.0 / c.H);
    const int cnt = nl.count[i];
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
        double xj[D], d2, r;
        load_vec<D>(s.x, j, xj);
        pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
        T W, fg;
        wendland<T>(T(r), T(c.H), inv_H, T(c.Wc), W, fg);
        dacc.add(s, ax, j, xi, xj, vi, gri, rhoi, d2, T(r), fg);
    }
    dacc.store(c, s, i);
}

void launch_surface1(int dim, int prec, const Consts& c, DevState s, Aux a, const int32_t* part_id, NbrList nl,
                     int64_t n, int r_from_x, double* norms_tmp, int with_dissipation, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) {
        if (prec == FJSPH_PREC_MIXED) hipLaunchKernelGGL((k_surface1<2, float>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp, with_dissipation);
        else hipLaunchKernelGGL((k_surface1<2, double>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp, with_dissipation);
    } else {
        if (prec == FJSPH_PREC_MIXED) hipLaunchKernelGGL((k_surface1<3, float>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp, with_dissipation);
        else hipLaunchKernelGGL((k_surface1<3, double>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp, with_dissipation);
    }
}

void launch_dissipation(int dim, int prec, const Consts& c, DevState s, Aux a, const int32_t* part_id, NbrList nl,
                        int64_t n, int r_from_x, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) {
        if (prec == FJSPH_PREC_MIXED) hipLaunchKernelGGL((k_dissipation<2, float>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x);
        else hipLaunchKernelGGL((k_dissipation<2, double>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x);
    } else {
        if (prec == FJSPH_PREC_MIXED) hipLaunchKernelGGL((k_dissipation<3, float>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x);
        else hipLaunchKernelGGL((k_dissipation<3, double>), g, b, 0, st, c, s, a, part_id, nl, n, r_from_x);
    }
}

/* Curvature, Gissler occlusion and surfzone (src/Geometry.cpp:145-277). */
template <int D>
__global__ __launch_bounds__(TPB) void k_surface2(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                  NbrList nl, int64_t n, int r_from_x,
                                                  const double* __restrict__ norms_tmp)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    double xi[D], ni[D], Vd[D];
    load_vec<D>(s.x, i, xi);
    load_vec<D>(norms_tmp, i, ni);
    const int32_t bi = s.b[i];
    const bool occl = (c.aero_case == FJSPH_GISSLER) && bi == FJSPH_FREE;
#pragma unroll
    for (int a = 0; a < D; ++a) Vd[a] = 0.0;
    if (s.cellID[i] != -1 && occl) {
#pragma unroll
        for (int a = 0; a < D; ++a)
            Vd[a] = (c.aero_source == FJSPH_MESH_INFL ? s.cellV[i * D + a] : c.v_inf[a]) - s.v[i * D + a];
    }
    const double Vdn = sqrt(dotd<D>(Vd, Vd));
    const bool ni_nz = dotd<D>(ni, ni) > 0.0;
    double Li[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b2 = 0; b2 < D; ++b2) Li[a][b2] = s.L[i * D * D + a * D + b2];
    const double inv_H = 1.0 / c.H;
    double curve = 0.0, qmax = 0.0;
    int zone = s.surf[i] == 1 ? 1 : 0;
    const int cnt = nl.count[i];
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
        double xj[D], d2, r, nj[D];
        load_vec<D>(s.x, j, xj);
        pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
        double Rij[D];
#pragma unroll
        for (int a = 0; a < D; ++a) Rij[a] = xj[a] - xi[a];
        load_vec<D>(norms_tmp, j, nj);
        if (ni_nz && dotd<D>(nj, nj) > 0.0) {
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double la = 0.0;
#pragma unroll
                for (int b2 = 0; b2 < D; ++b2) la += Li[a][b2] * (nj[b2] - ni[b2]);
                t += la * (fg * Rij[a]);
            }
            curve += ax.vol[j] * t;
        }
        if (occl) {
            /* frac = -Rij.Vd / (|Vd| r); the max is taken over -Rij.Vd / r
             * and divided by |Vd| once */
            const double q = -dotd<D>(Rij, Vd) / r;
            if (q > qmax) qmax = q;
        }
        if (c.ale && s.surf[j] == 1) zone = 1;
    }
    const double wo = (occl && Vdn > 0.0) ? qmax / Vdn : 0.0;
    s.woccl[i] = s.lam_nb[i] < c.lam_cutoff ? fmax(0.0, fmin(wo, 1.0)) : 1.0;
    double nh[D];
    normalize<D>(ni, nh);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        s.norm[i * D + a] = ni[a];
        ax.nhat[i * D + a] = nh[a];
    }
    s.curve[i] = curve;
    s.norm_curve[i] = c.dx * curve;
    s.pDist[i] = s.lam[i];
    if (c.ale) s.surfzone[i] = zone;
}

void launch_surface2(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, NbrList nl, int64_t n,
                     int r_from_x, const double* norms_tmp, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_surface2<2>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp);
    else hipLaunchKernelGGL(k_surface2<3>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x, norms_tmp);
}

/* ------------------------------------------------------------ particle_shift */

/* src/Shifting.cpp:189-290.  max_j acos(n_i.n_j) is taken as acos(min_j
 * n_i.n_j) over the dot products acos accepts ([-1, 1]); NaN angles never
 * win the reference's max, and acos is decreasing, so the result is the
 * same.  max_j |v_j - v_i| is taken on squared norms. */
template <int D>
__global__ __launch_bounds__(TPB) void k_shift(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                               NbrList nl, int64_t n, int r_from_x)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    const double lam_nb = s.lam_nb[i];
    if (lam_nb < 0.55 || s.b[i] == FJSPH_BUFFER) {
#pragma unroll
        for (int a = 0; a < D; ++a) s.vPert[i * D + a] = 0.0;
        return;
    }
    double xi[D], vi[D], nhat[D];
    load_vec<D>(s.x, i, xi);
    load_vec<D>(s.v, i, vi);
    load_vec<D>(ax.nhat, i, nhat);
    const double inv_H = 1.0 / c.H;
    const double inv_Wdx = 1.0 / c.W_dx;
    double du[D];
#pragma unroll
    for (int a = 0; a < D; ++a) du[a] = 0.0;
    double maxU2 = 0.0, dmin = 2.0;
    const int cnt = nl.count[i];
    for (int k = 0; k < cnt; ++k) {
        const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
        double xj[D], d2, r;
        load_vec<D>(s.x, j, xj);
        pair_dist<D>(r_from_x, xi, xj, nl, k, i, d2, r);
        double W, fg;
        wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
        const double w = W * inv_Wdx;
        const double w2 = w * w;
        const double f = (1.0 + 0.2 * (w2 * w2)) * fg * ax.vol[j];
        double dv2 = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            du[a] += f * (xj[a] - xi[a]);
            const double dv = s.v[(int64_t)j * D + a] - vi[a];
            dv2 += dv * dv;
        }
        if (s.b[j] > FJSPH_PISTON) {
            double nj[D];
            load_vec<D>(ax.nhat, j, nj);
            const double d = dotd<D>(nhat, nj);
            if (d >= -1.0 && d <= 1.0 && d < dmin) dmin = d;
        }
        if (dv2 > maxU2) maxU2 = dv2;
    }
    const double wo = dmin <= 1.0 ? acos(dmin) : 0.0;
    const double fac = -2.0 * c.H * sqrt(dotd<D>(vi, vi));
#pragma unroll
    for (int a = 0; a < D; ++a) du[a] *= fac;
    double dun[D];
    normalize<D>(du, dun);
    const double mag = fmin(sqrt(dotd<D>(du, du)), fmin(sqrt(maxU2) / 2.0, c.max_shift_vel));
#pragma unroll
    for (int a = 0; a < D; ++a) du[a] = mag * dun[a];
    double out[D];
    if (s.surfzone[i] == 0 && lam_nb > 0.55) {
#pragma unroll
        for (int a = 0; a < D; ++a) out[a] = du[a];
    } else if (wo < kPi / 12.0) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            double t = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) t += (((a == b2) ? 1.0 : 0.0) - nhat[a] * nhat[b2]) * du[b2];
            out[a] = t;
        }
    } else {
#pragma unroll
        for (int a = 0; a < D; ++a) out[a] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) s.vPert[i * D + a] = out[a];
}

void launch_shift(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, NbrList nl, int64_t n,
                  int r_from_x, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_shift<2>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x);
    else hipLaunchKernelGGL(k_shift<3>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, a, part_id, nl, n, r_from_x);
}

/* ------------------------------------------------------- boundary solvers */

/* Block velocity + no-slip (src/Newmark_Beta.cpp:69-102, src/Resid.cpp:162-186). */
template <int D>
__global__ __launch_bounds__(TPB) void k_bnd_vel(Consts c, DevState s, const int32_t* __restrict__ blk, DevBlocks bl,
                                                 int32_t block, NbrList nl, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || blk[i] != block) return;
    double v[D];
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = bl.vel[block * 3 + a];
    if (bl.no_slip[block]) {
        const double inv_H = 1.0 / c.H;
        double xi[D], vs[D];
        load_vec<D>(s.x, i, xi);
#pragma unroll
        for (int a = 0; a < D; ++a) vs[a] = 0.0;
        double ks = 0.0;
        const int cnt = nl.count[i];
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
            if (s.b[j] <= FJSPH_PISTON) continue;
            double xj[D], d2, r;
            load_vec<D>(s.x, j, xj);
            pair_dist<D>(0, xi, xj, nl, k, i, d2, r);
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            ks += W;
#pragma unroll
            for (int a = 0; a < D; ++a) vs[a] += s.v[(int64_t)j * D + a] * W;
        }
        if (ks > 0.0) {
#pragma unroll
            for (int a = 0; a < D; ++a) v[a] = 2.0 * v[a] - vs[a] / ks;
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) s.v[i * D + a] = v[a];
}

/* Boundary solver of one block (src/Resid.cpp:21-160). The list distance is
 * the frozen one, as in the reference. */
template <int D>
__global__ __launch_bounds__(TPB) void k_bnd_solve(Consts c, DevState s, Aux ax, const int32_t* __restrict__ blk,
                                                   DevBlocks bl, int32_t block, NbrList nl, int64_t n,
                                                   int32_t* near_inlet)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || blk[i] != block) return;
    const int solver = bl.solver[block];
    const double inv_H = 1.0 / c.H;
    const int cnt = nl.count[i];
    double xi[D];
    load_vec<D>(s.x, i, xi);
    if (solver == FJSPH_DBC) {
#pragma unroll
        for (int a = 0; a < D; ++a) s.acc[i * D + a] = 0.0;
    } else if (solver == FJSPH_PRESSURE_G) {
        int near = 0;
        double ks = 0.0, pk = 0.0, ak[D];
#pragma unroll
        for (int a = 0; a < D; ++a) ak[a] = 0.0;
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
            if (s.b[j] <= FJSPH_PISTON) continue;
            double xj[D], d2, r;
            load_vec<D>(s.x, j, xj);
            pair_dist<D>(0, xi, xj, nl, k, i, d2, r);
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            const double kern = ax.vol[j] * W;
            ks += kern;
            pk += s.p[j] * kern;
            const double kr = kern * s.rho[j];
#pragma unroll
            for (int a = 0; a < D; ++a) ak[a] += kr * (xi[a] - xj[a]);
            if (s.surfzone[j]) near = 1;
        }
        double pres = 0.0;
        if (ks > 0.0) {
            double ga = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) ga += (c.grav[a] - s.acc[i * D + a]) * ak[a];
            const double p = (pk + ga) / ks;
            pres = near ? fmax(0.0, p) : p;
        }
        s.p[i] = pres;
        s.rho[i] = eos_density(c, pres);
    } else if (solver == FJSPH_GHOST) {
        int ni = 1;
        double rr = 0.0, vi[D];
        load_vec<D>(s.v, i, vi);
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = nl.idx[(int64_t)k * nl.npad + i];
            const int32_t bj = s.b[j];
            if (bj == FJSPH_PIPE || bj == FJSPH_FREE) ni = 0;
            double xj[D], d2, r;
            load_vec<D>(s.x, j, xj);
            pair_dist<D>(0, xi, xj, nl, k, i, d2, r);
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) t += (s.v[(int64_t)j * D + a] - vi[a]) * (fg * (xj[a] - xi[a]));
            rr -= ax.vol[j] * t;
        }
        near_inlet[i] = ni;
        s.Rrho[i] = rr * s.rho[i];
    }
}

void launch_boundary(int dim, const Consts& c, DevState s, Aux a, const int32_t* blk, DevBlocks bl, int32_t block,
                     NbrList nl, int64_t n, int32_t* near_inlet, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) {
        hipLaunchKernelGGL(k_bnd_vel<2>, g, b, 0, st, c, s, blk, bl, block, nl, n);
        hipLaunchKernelGGL(k_bnd_solve<2>, g, b, 0, st, c, s, a, blk, bl, block, nl, n, near_inlet);
    } else {
        hipLaunchKernelGGL(k_bnd_vel<3>, g, b, 0, st, c, s, blk, bl, block, nl, n);
        hipLaunchKernelGGL(k_bnd_solve<3>, g, b, 0, st, c, s, a, blk, bl, block, nl, n, near_inlet);
    }
}

/* ----------------------------------------------------- force + Newmark */

/* Gissler drag (src/Aero.h:37-98) in FP64 for one particle. */
template <int D>
__device__ void gissler(const Consts& c, const double* Vd, double rho, double press, double mass, double lam,
                        double nneigh, double woccl, double* out)
{
    const double Vn = sqrt(dotd<D>(Vd, Vd));
    const double Re = 2.0 * rho * Vn * c.L / c.mu_g;
    const double frac2 = c.use_lam ? fmin(c.interp_fac * lam, 1.0) : fmin(c.interp_fac * nneigh * c.i_n_full, 1.0);
    const double frac1 = 1.0 - frac2;
    const double Cds = (1.0 + 0.197 * pow(Re, 0.63) + 2.6e-04 * pow(Re, 1.38)) * (24.0 / (Re + 0.00001));
    double Cdl, Adrop;
    if (c.use_TAB_def) {
        double ymax = dotd<D>(Vd, Vd) * c.ycoef;
        if (ymax > 1.0) ymax = 1.0;
        Cdl = Cds * (1 + 2.632 * ymax);
        if (D == 3) {
            const double t = c.L + c.Cb * c.L * ymax;
            Adrop = kPi * t * t;
        } else {
            Adrop = c.A_sphere + 2 * (c.Cb * c.L * ymax);
        }
    } else {
        Cdl = Cds;
        Adrop = c.A_sphere;
    }
    const double Cdi = frac1 * Cdl + frac2;
    const double Ai = (1.0 - woccl) * (frac1 * Adrop + frac2 * c.A_plate);
    const double f = 0.5 * Vn / (c.sos * c.sos) * c.gamma_g * press * Cdi * Ai / mass;
#pragma unroll
    for (int a = 0; a < D; ++a) out[a] = f * Vd[a];
}

struct NextState {
    double *x, *v, *rho, *p;
};

/* get_acc_and_Rrho_on_i (src/Resid.cpp:243-418) for fluid slots, fused with
 * the Newmark-Beta update of Do_NB_Iter (src/Newmark_Beta.cpp:137-240) when
 * do_update.  Pair terms are formed in T; every sum is accumulated in FP64.
 * The pressure, viscous and surface-tension terms of a pair are combined
 * before accumulation: acc_i = sum_j (-m_j (P_i + P_j) g + visc + ST / m_i). */
template <int D, typename T>
__global__ __launch_bounds__(TPB) void k_force(Consts c, DevState s, Aux ax, DevState pn, NextState nx, int do_update,
                                               const int32_t* __restrict__ part_id, const int32_t* __restrict__ blk,
                                               DevBlocks bl, const int32_t* __restrict__ near_inlet, NbrList nl,
                                               int64_t n, double dt, const DevScalars* sc, double* partial)
{
    __shared__ double lds[NWAVE];
    constexpr bool F32 = sizeof(T) == 4;
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    double resid = 0.0;
    if (i < n) {
        const bool fluid = part_id[i] >= c.n_bound;
        double acc_new[D], Rrho_new = 0.0;
        if (fluid) {
            double xi[D], vi[D];
            load_vec<D>(s.x, i, xi);
            load_vec<D>(s.v, i, vi);
            const double rhoi = s.rho[i], mi = s.m[i];
            const int32_t bi = s.b[i];
            const double npd = sc->npd;
            const T ste = T(-(0.5 * c.sig / c.lam_st) / (npd * npd));
            const T pi3o4 = T(3.0 * kPi / 4.0);
            const T inv_H = T(1.0 / c.H), H = T(c.H), Wc = T(c.Wc);
            const T eps2 = T(0.001 * c.H_sq);
            const T Pi = T(ax.P[i]);
            const T iri = T(ax.invrho[i]);
            const T nu = T(c.nu);
            const T inv_mi = T(1.0 / mi);
            const bool st_on = (!c.ale) || (s.surfzone[i] == 1);
            const bool bound_i = bi == FJSPH_BOUND;
            const T contang_fac = T(1.0 + 0.5 * cos(0.5 * kPi * 7.0 / 9.0));
            double acc[D];
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] = 0.0;
            double Rrho_ = 0.0, Rrhoc_ = 0.0;
            const int cnt = nl.count[i];
            if (s.cellID[i] != -1) {
                double Vd[D], aero[D];
#pragma unroll
                for (int a = 0; a < D; ++a) Vd[a] = s.cellV[i * D + a] - vi[a];
                if (c.aero_case == FJSPH_GISSLER) {
                    gissler<D>(c, Vd, s.cellRho[i], s.cellP[i], mi, s.lam_nb[i], double(cnt + 1), s.woccl[i], aero);
                } else {
#pragma unroll
                    for (int a = 0; a < D; ++a) aero[a] = 0.0;
                }
#pragma unroll
                for (int a = 0; a < D; ++a) { acc[a] += aero[a]; s.Af[i * D + a] = aero[a]; }
            }
            T vpi[D];
            if (c.ale) {
#pragma unroll
                for (int a = 0; a < D; ++a) vpi[a] = T(s.vPert[i * D + a]);
            }
            for (int k = 0; k < cnt; ++k) {
                const int64_t e = (int64_t)k * nl.npad + i;
                const int32_t j = nl.idx[e];
                T r, rr;
                if (F32) {
                    r = T(((const float*)nl.r)[e]);
                    rr = r * r;
                } else {
                    const double d2 = ((const double*)nl.r)[e];
                    rr = T(d2);
                    r = T(sqrt(d2));
                }
                T Rji[D], Vji[D];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    Rji[a] = T(s.x[(int64_t)j * D + a] - xi[a]);
                    Vji[a] = T(s.v[(int64_t)j * D + a] - vi[a]);
                }
                T mj, Pj, irj, volj;
                bool bound_j;
                if (F32) {
                    const float4 q = ax.fq[j];
                    bound_j = q.x < 0.f;
                    mj = T(fabsf(q.x));
                    Pj = T(q.y);
                    irj = T(q.z);
                    volj = T(q.w);
                } else {
                    mj = T(s.m[j]);
                    Pj = T(ax.P[j]);
                    irj = T(ax.invrho[j]);
                    volj = T(ax.vol[j]);
                    bound_j = s.b[j] == FJSPH_BOUND;
                }
                const T idist2 = fast_rcp<T>(rr + eps2);
                T W, fg;
                wendland<T>(r, H, inv_H, Wc, W, fg);
                T g[D];
#pragma unroll
                for (int a = 0; a < D; ++a) g[a] = fg * Rji[a];
                const T pc = mj * (Pi + Pj);
                T rdotg = Rji[0] * g[0];
#pragma unroll
                for (int a = 1; a < D; ++a) rdotg += Rji[a] * g[a];
                const T vf = mj * nu * (iri + irj) * rdotg * idist2;
                T sf = T(0);
                if (st_on) {
                    const T fac = (bound_i || bound_j) ? contang_fac : T(1);
                    sf = ste * fac * cos(pi3o4 * r * inv_H) * fast_rcp<T>(r) * inv_mi;
                }
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += double(vf * Vji[a] + sf * Rji[a] - pc * g[a]);
                if (c.ale) {
                    T vpj[D], rhoj;
                    if (F32) {
                        const float4 q = ax.fale[j];
                        vpj[0] = q.x;
                        if (D > 1) vpj[1 % D] = q.y;
                        if (D > 2) vpj[2 % D] = q.z;
                        rhoj = q.w;
                    } else {
#pragma unroll
                        for (int a = 0; a < D; ++a) vpj[a] = T(s.vPert[(int64_t)j * D + a]);
                        rhoj = T(s.rho[j]);
                    }
                    T aj = T(0), ai = T(0), cd = T(0), cont = T(0), c2 = T(0);
                    const T rhoit = T(rhoi);
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        aj += vpj[a] * g[a];
                        ai += vpi[a] * g[a];
                        cd += (vpj[a] - vpi[a]) * g[a];
                        cont += (Vji[a] + (vpj[a] - vpi[a])) * g[a];
                        c2 += (rhoj * vpj[a] + rhoit * vpi[a]) * g[a];
                    }
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        const T vja = T(s.v[(int64_t)j * D + a]), via = T(vi[a]);
                        acc[a] += double((vja * aj + via * ai - via * cd) * volj);
                    }
                    Rrho_ -= double(cont * volj);
                    Rrhoc_ += double(c2 * volj);
                } else {
                    T vg = Vji[0] * g[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) vg += Vji[a] * g[a];
                    Rrho_ -= double(volj * vg);
                }
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                acc_new[a] = acc[a] + s.aVisc[i * D + a] + c.grav[a];
                s.acc[i * D + a] = acc_new[a];
            }
            Rrho_new = c.ale ? (Rrho_ * rhoi + Rrhoc_ + s.deltaD[i]) : (Rrho_ * rhoi + s.deltaD[i]);
            s.Rrho[i] = Rrho_new;
        }
        if (do_update) {
            /* Newmark-Beta update (beta = 1/4, gamma = 1/2) */
            const double g1 = 0.5, g2 = 0.5, b1 = 0.25, b2 = 0.25;
            const double dt2 = dt * dt;
            double xo[D], vo[D], rho_o = s.rho[i], p_o = s.p[i];
            load_vec<D>(s.x, i, xo);
            load_vec<D>(s.v, i, vo);
            const int32_t bi = s.b[i];
            if (fluid) {
                if (bi > FJSPH_BUFFER && bi != FJSPH_OUTLET) {
                    double xn[D], vn[D], an[D];
                    load_vec<D>(pn.x, i, xn);
                    load_vec<D>(pn.v, i, vn);
                    load_vec<D>(pn.acc, i, an);
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        double xnew;
                        if (c.ale)
                            xnew = xn[a] + dt * (vn[a] + s.vPert[i * D + a]) + dt2 * (b1 * acc_new[a] + b2 * an[a]);
                        else
                            xnew = xn[a] + dt * vn[a] + dt2 * (b2 * an[a] + b1 * acc_new[a]);
                        const double d = xnew - xo[a];
                        resid += d * d;
                        xo[a] = xnew;
                        vo[a] = vn[a] + dt * (g1 * acc_new[a] + g2 * an[a]);
                    }
                    rho_o = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dt * (g1 * Rrho_new + g2 * pn.Rrho[i])));
                    p_o = eos_pressure(c, rho_o);
                } else if (bi == FJSPH_OUTLET) {
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        const double xnew = pn.x[i * D + a] + dt * vo[a];
                        const double d = xnew - xo[a];
                        resid += d * d;
                        xo[a] = xnew;
                    }
                }
            } else {
                const int bk = blk[i];
                const int solver = bk >= 0 ? bl.solver[bk] : -1;
                if (solver == FJSPH_DBC) {
                    rho_o = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dt * (g1 * s.Rrho[i] + g2 * pn.Rrho[i])));
                    p_o = eos_pressure(c, rho_o);
                } else if (solver == FJSPH_GHOST) {
                    const double lo = near_inlet[i] ? c.rho0 : c.rho_min;
                    rho_o = fmax(lo, fmin(c.rho_max, pn.rho[i] + dt * (g1 * s.Rrho[i] + g2 * pn.Rrho[i])));
                    p_o = eos_pressure(c, rho_o);
                    if (near_inlet[i]) s.Rrho[i] = fmax(0.0, s.Rrho[i]);
                }
            }
#pragma unroll
            for (int a = 0; a < D; ++a) { nx.x[i * D + a] = xo[a]; nx.v[i * D + a] = vo[a]; }
            nx.rho[i] = rho_o;
            nx.p[i] = p_o;
        }
    }
    if (do_update) {
        const double bs = block_reduce(resid, OpSum(), lds);
        if (threadIdx.x == 0) partial[blockIdx.x] = bs;
    }
}

void launch_force(int dim, int prec, const Consts& c, DevState s, Aux a, DevState pn, double* next_x,
                  double* next_v, double* next_rho, double* next_p, const int32_t* part_id, const int32_t* blk,
                  DevBlocks bl, const int32_t* near_inlet, NbrList nl, int64_t n, double dt, const DevScalars* sc,
                  double* partial, int nparts, hipStream_t st)
{
    NextState nx{next_x, next_v, next_rho, next_p};
    const int upd = next_x != nullptr;
    const dim3 g(nparts), b(TPB);
    if (dim == 2) {
        if (prec == FJSPH_PREC_MIXED)
            hipLaunchKernelGGL((k_force<2, float>), g, b, 0, st, c, s, a, pn, nx, upd, part_id, blk, bl, near_inlet, nl, n, dt, sc, partial);
        else
            hipLaunchKernelGGL((k_force<2, double>), g, b, 0, st, c, s, a, pn, nx, upd, part_id, blk, bl, near_inlet, nl, n, dt, sc, partial);
    } else {
        if (prec == FJSPH_PREC_MIXED)
            hipLaunchKernelGGL((k_force<3, float>), g, b, 0, st, c, s, a, pn, nx, upd, part_id, blk, bl, near_inlet, nl, n, dt, sc, partial);
        else
            hipLaunchKernelGGL((k_force<3, double>), g, b, 0, st, c, s, a, pn, nx, upd, part_id, blk, bl, near_inlet, nl, n, dt, sc, partial);
    }
}

/* ------------------------------------------------------ find_timestep */

/* Per fluid particle: |acc|, |Af|, |Rrho|, |rho - rho0|, |v|, |vPert| (max)
 * and the surface-tension limit sqrt(rho dx^2 / (2 pi sig |curve|)) (min,
 * skipped where the denominator is zero). */
template <int D>
__global__ __launch_bounds__(TPB) void k_dt_reduce(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                   int64_t n, double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    const double eps = 2.220446049250313e-16;
    double v[7] = {eps, eps, eps, eps, eps, eps, 9999999.0};
    if (i < n && part_id[i] >= c.n_bound) {
        double t[D];
        load_vec<D>(s.acc, i, t);
        v[0] = sqrt(dotd<D>(t, t));
        load_vec<D>(s.Af, i, t);
        v[1] = sqrt(dotd<D>(t, t));
        v[2] = fabs(s.Rrho[i]);
        v[3] = fabs(s.rho[i] - c.rho0);
        load_vec<D>(s.v, i, t);
        v[4] = sqrt(dotd<D>(t, t));
        load_vec<D>(s.vPert, i, t);
        v[5] = sqrt(dotd<D>(t, t));
        const double den = 2.0 * kPi * c.sig * fabs(s.curve[i]);
        if (den != 0.0) v[6] = sqrt(s.rho[i] * c.dx * c.dx / den);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double r = block_reduce(v[k], OpMax(), lds);
        if (threadIdx.x == 0) partial[blockIdx.x * 7 + k] = r;
    }
    const double r = block_reduce(v[6], OpMin(), lds);
    if (threadIdx.x == 0) partial[blockIdx.x * 7 + 6] = r;
}

__global__ __launch_bounds__(TPB) void k_dt_final(const double* partial, int nparts, double* out)
{
    __shared__ double lds[NWAVE];
    for (int k = 0; k < 7; ++k) {
        double v = partial[k];
        for (int b = threadIdx.x; b < nparts; b += TPB) {
            const double w = partial[b * 7 + k];
            v = (k < 6) ? fmax(v, w) : fmin(v, w);
        }
        const double r = (k < 6) ? block_reduce(v, OpMax(), lds) : block_reduce(v, OpMin(), lds);
        if (threadIdx.x == 0) out[k] = r;
    }
}

void launch_timestep_reduce(int dim, const Consts& c, DevState s, const int32_t* part_id, int64_t n,
                            double* partial, int nparts, double* out, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_dt_reduce<2>, dim3(nparts), dim3(TPB), 0, st, c, s, part_id, n, partial);
    else hipLaunchKernelGGL(k_dt_reduce<3>, dim3(nparts), dim3(TPB), 0, st, c, s, part_id, n, partial);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(TPB), 0, st, partial, nparts, out);
}

__global__ __launch_bounds__(TPB) void k_fix_rho(DevState s, const int32_t* __restrict__ part_id, int64_t nb,
                                                 int64_t n, double rho0)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < nb) return;
    if (s.rho[i] < 0.0001) s.rho[i] = rho0;
}

void launch_fix_rho(DevState s, const int32_t* part_id, int64_t nb, int64_t n, double rho0, hipStream_t st)
{
    hipLaunchKernelGGL(k_fix_rho, dim3(grid_of(n)), dim3(TPB), 0, st, s, part_id, nb, n, rho0);
}

}  // namespace fj
