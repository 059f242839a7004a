/*
 * gfx950 kernels of the WCSPH step engine: cell binning, reorder,
 * per-particle derived quantities and reductions.
 *
 * One lane per particle slot; slots are sorted by cell so a wavefront's
 * neighbours sit in a few contiguous strips and the neighbour gathers hit L2.
 * Pair terms never scatter: every pass is owner-computes over the particle's
 * own list, so results do not depend on launch order.  Reductions are
 * deterministic (fixed-order wave / block / two-stage trees).
 *
 * Reference correspondence (file:line):
 *   k_cells/k_scan/k_scatter/k_sort_cells  cell binning of update_neighbours, src/Neighbours.cpp:7-47
 *   k_dt_reduce    find_timestep reduction   src/Integration.cpp:384-403
 * The pair passes live in pairs.hip.
 */
#include "device_common.hpp"

namespace fj {

int reduction_parts(int64_t n) { return (int)grid_of(n); }

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
    /* independent simulations (grp != nullptr): each group is binned from
     * its own lower corner glo[3 g..] into its own cell layers along the
     * last axis, gl0[2 g] (first) and gl0[2 g + 1] (count), so the groups
     * sit in disjoint layers in group order */
    const int32_t* grp;
    const double* glo;
    const int32_t* gl0;
};

/* Runs of equal consecutive values in a wave (every lane takes part):
 * the lane's run head and, for a head, the run length.  The particles are
 * in cell order after a permutation and nearly so between permutations, so
 * one atomic per run replaces one per particle. */
__device__ __forceinline__ void wave_runs(int32_t v, int lane, int& head, int& len)
{
    const int32_t prev = __shfl_up(v, 1, 64);
    const bool is_head = lane == 0 || prev != v;
    const uint64_t heads = __ballot(is_head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    head = 63 - __clzll(heads & upto);
    const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1));
    len = after ? __ffsll((unsigned long long)after) : 64 - lane;
}

template <int D>
__global__ __launch_bounds__(TPB) void k_cells(const double* __restrict__ x, int64_t n, Grid3 g, int32_t* cell,
                                               int32_t* cell_count)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    int32_t lin = -1 - lane; /* lanes past n: distinct values, runs of their own */
    if (valid) {
        int32_t ci[3] = {0, 0, 0};
        if (g.grp) {
            /* a particle outside its group's box (after a rollback) is
             * clamped into the group's own cells: clamping keeps
             * neighbouring particles in neighbouring cells */
            const int32_t gi = g.grp[i];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double f = floor((x[i * D + k] - g.glo[gi * 3 + k]) * g.inv_cs);
                const double top = (k == D - 1) ? double(g.gl0[2 * gi + 1] - 1) : double(g.dims[k] - 1);
                ci[k] = (int32_t)fmin(fmax(f, 0.0), top) + (k == D - 1 ? g.gl0[2 * gi] : 0);
            }
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                double f = floor((x[i * D + k] - g.lo[k]) * g.inv_cs);
                int32_t c = (int32_t)fmin(fmax(f, 0.0), double(g.dims[k] - 1));
                ci[k] = c;
            }
        }
        lin = (D == 3) ? ((ci[2] * g.dims[1] + ci[1]) * g.dims[0] + ci[0]) : (ci[1] * g.dims[0] + ci[0]);
        cell[i] = lin;
    }
    int head, len;
    wave_runs(lin, lane, head, len);
    if (valid && head == lane) atomicAdd(&cell_count[lin], len);
}

void launch_cells(int dim, const double* x, int64_t n, const double* lo, double inv_cs, const int32_t* dims,
                  int32_t* cell, int32_t* cell_count, hipStream_t st, const int32_t* grp, const double* glo,
                  const int32_t* gl0)
{
    Grid3 g;
    for (int k = 0; k < 3; ++k) { g.lo[k] = k < dim ? lo[k] : 0.0; g.dims[k] = dims[k]; }
    g.inv_cs = inv_cs;
    g.grp = grp;
    g.glo = glo;
    g.gl0 = gl0;
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
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const int32_t c = valid ? cell[i] : -1 - lane;
    int head, len;
    wave_runs(c, lane, head, len);
    int32_t base = 0;
    if (valid && head == lane) base = atomicAdd(&fill[c], len);
    base = __shfl(base, head, 64);
    if (valid) order[start[c] + base + (lane - head)] = (int32_t)i;
}

void launch_scatter(const int32_t* cell, int64_t n, const int32_t* start, int32_t* fill, int32_t* order,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(TPB), 0, st, cell, n, start, fill, order);
}

/* Sort each cell's slot list ascending so list order is deterministic: a
 * workgroup per cell at a time, segment staged in LDS, each lane writes its
 * elements at their rank.  Segments longer than the LDS tile fall back to an
 * insertion sort in one lane. */
constexpr int SORT_TILE = 1024;

__global__ __launch_bounds__(64) void k_sort_cells(const int32_t* __restrict__ start, int64_t ncell, int32_t* order)
{
    __shared__ int32_t seg[SORT_TILE];
    /* cells dealt to the workgroups in turn (sparse grids are mostly empty
     * cells, which a workgroup each would only launch and retire) */
    for (int64_t c = blockIdx.x; c < ncell; c += gridDim.x) {
        const int32_t a = start[c], b = start[c + 1];
        const int32_t k = b - a;
        if (k <= 1) continue;
        if (k > SORT_TILE) {
            if (threadIdx.x == 0) {
                for (int32_t i = a + 1; i < b; ++i) {
                    const int32_t v = order[i];
                    int32_t j = i - 1;
                    while (j >= a && order[j] > v) { order[j + 1] = order[j]; --j; }
                    order[j + 1] = v;
                }
            }
            continue;
        }
        __syncthreads(); /* seg is read by the previous cell's ranking */
        for (int t = threadIdx.x; t < k; t += 64) seg[t] = order[a + t];
        __syncthreads();
        for (int t = threadIdx.x; t < k; t += 64) {
            const int32_t v = seg[t];
            int rank = 0;
            for (int u = 0; u < k; ++u) rank += seg[u] < v;
            order[a + rank] = v;
        }
    }
}

void launch_sort_cells(const int32_t* start, int64_t ncell, int32_t* order, hipStream_t st)
{
    const int64_t g = std::min<int64_t>(std::max<int64_t>(ncell, 1), 65536);
    hipLaunchKernelGGL(k_sort_cells, dim3((unsigned)g), dim3(64), 0, st, start, ncell, order);
}

/* Per-particle derived quantities (see Aux). */
template <int D>
__global__ __launch_bounds__(TPB) void k_aux(DevState s, Aux a, int64_t n, int ale, int mixed, float bgam)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    const double rho = s.rho[i], m = s.m[i], p = s.p[i];
    const double vol = m / rho, P = p / (rho * rho), ir = 1.0 / rho;
    a.vol[i] = vol;
    a.P[i] = P;
    a.invrho[i] = ir;
    if (mixed) {
        const int32_t b = s.b[i];
        const float ms = b == FJSPH_BOUND ? -float(m) : float(m);
        a.fq[i] = make_float4(ms, float(P), float(ir), float(vol));
        a.fs2[i] = make_float4(float(vol), __builtin_amdgcn_sqrtf(bgam * float(ir)),
                               b > FJSPH_PISTON ? float(m) : -float(m), 0.f);
        if (ale) {
            float vp[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < D; ++k) vp[k] = float(s.vPert[i * D + k]);
            a.fale[i] = make_float4(vp[0], vp[1], vp[2], float(rho));
        }
    }
}

void launch_aux(int dim, DevState s, Aux a, int64_t n, int ale, int mixed, double bgam, hipStream_t st)
{
    const float bg = float(bgam);
    if (dim == 2) hipLaunchKernelGGL(k_aux<2>, dim3(grid_of(n)), dim3(TPB), 0, st, s, a, n, ale, mixed, bg);
    else hipLaunchKernelGGL(k_aux<3>, dim3(grid_of(n)), dim3(TPB), 0, st, s, a, n, ale, mixed, bg);
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

/* Sum `width` columns of per-block partials in a fixed order, times scale. */
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

void launch_sum_partials_scaled(const double* partial, int nparts, int width, double scale, double* out,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(TPB), 0, st, partial, nparts, width, scale, out);
}

/* ------------------------------------------------------ find_timestep */

/* Per fluid particle: |acc|, |Af|, |Rrho|, |rho - rho0|, |v|, |vPert| (max)
 * and the surface-tension limit sqrt(rho dx^2 / (2 pi sig |curve|)) (min,
 * skipped where the denominator is zero). */
template <int D>
__device__ __forceinline__ void dt_terms(const Consts& c, const DevState& s, int64_t i, double* v)
{
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

template <int D>
__global__ __launch_bounds__(TPB) void k_dt_reduce(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                   int64_t n, double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    const double eps = 2.220446049250313e-16;
    double v[7] = {eps, eps, eps, eps, eps, eps, 9999999.0};
    if (i < n && part_id[i] >= c.n_bound) dt_terms<D>(c, s, i, v);
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

/* ------------------------------------------ independent simulations */

/* Bounding box of each group's slots [gstart[g], gstart[g+1]) (slots are
 * group-major): one block per group, lo then hi per axis in out[6 g..]. */
template <int D>
__global__ __launch_bounds__(TPB) void k_group_bbox(const double* __restrict__ x, const int64_t* __restrict__ gstart,
                                                    double* out)
{
    __shared__ double lds[NWAVE];
    const int64_t g = blockIdx.x, a = gstart[g], e = gstart[g + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double lo = 1e300, hi = -1e300;
        for (int64_t i = a + threadIdx.x; i < e; i += TPB) {
            const double v = x[i * D + k];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        lo = block_reduce(lo, OpMin(), lds);
        hi = block_reduce(hi, OpMax(), lds);
        if (threadIdx.x == 0) {
            out[g * 6 + k] = lo;
            out[g * 6 + 3 + k] = hi;
        }
    }
}

/* find_timestep's reduction per group (slots [gstart[g], gstart[g+1]),
 * fluid only): out[8 g + k] as k_dt_final's out[k]. */
template <int D>
__global__ __launch_bounds__(TPB) void k_group_dt(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                  const int64_t* __restrict__ gstart, double* out)
{
    __shared__ double lds[NWAVE];
    const int64_t g = blockIdx.x, a = gstart[g], e = gstart[g + 1];
    const double eps = 2.220446049250313e-16;
    double r[7] = {eps, eps, eps, eps, eps, eps, 9999999.0};
    for (int64_t i = a + threadIdx.x; i < e; i += TPB) {
        if (part_id[i] < c.n_bound) continue;
        double v[7] = {eps, eps, eps, eps, eps, eps, 9999999.0};
        dt_terms<D>(c, s, i, v);
#pragma unroll
        for (int k = 0; k < 6; ++k) r[k] = fmax(r[k], v[k]);
        r[6] = fmin(r[6], v[6]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double m = block_reduce(r[k], OpMax(), lds);
        if (threadIdx.x == 0) out[g * 8 + k] = m;
    }
    const double m = block_reduce(r[6], OpMin(), lds);
    if (threadIdx.x == 0) out[g * 8 + 6] = m;
}

void launch_group_dt(int dim, const Consts& c, DevState s, const int32_t* part_id, const int64_t* gstart,
                     int32_t ngroups, double* out, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_group_dt<2>, dim3(ngroups), dim3(TPB), 0, st, c, s, part_id, gstart, out);
    else hipLaunchKernelGGL(k_group_dt<3>, dim3(ngroups), dim3(TPB), 0, st, c, s, part_id, gstart, out);
}

void launch_group_bbox(int dim, const double* x, const int64_t* gstart, int32_t ngroups, double* out, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_group_bbox<2>, dim3(ngroups), dim3(TPB), 0, st, x, gstart, out);
    else hipLaunchKernelGGL(k_group_bbox<3>, dim3(ngroups), dim3(TPB), 0, st, x, gstart, out);
}

/* out[g] = scale[g] * sum of vals[seg[g] .. seg[g+1]) in a fixed order
 * (scale may be null: 1), one block per segment. */
__global__ __launch_bounds__(TPB) void k_seg_sum(const double* __restrict__ vals, const int64_t* __restrict__ seg,
                                                 const double* __restrict__ scale, double* __restrict__ out)
{
    __shared__ double lds[NWAVE];
    const int64_t g = blockIdx.x, a = seg[g], e = seg[g + 1];
    double acc = 0.0;
    for (int64_t i = a + threadIdx.x; i < e; i += TPB) acc += vals[i];
    acc = block_reduce(acc, OpSum(), lds);
    if (threadIdx.x == 0) out[g] = scale ? acc * scale[g] : acc;
}

void launch_seg_sum(const double* vals, const int64_t* seg, int32_t nseg, const double* scale, double* out,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_seg_sum, dim3(nseg), dim3(TPB), 0, st, vals, seg, scale, out);
}

}  // namespace fj
