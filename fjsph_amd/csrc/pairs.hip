/*
 * Block-tiled pair passes of the WCSPH step.
 *
 * Work decomposition: a block is a chunk of C consecutive cells of one cell
 * row; its targets are the particles of those cells (positions [p0, p1) of
 * the cell-sorted order[] array, recorded as the target order vorder[]) and
 * its candidate union is the 3^(D-1) neighbour rows over cells
 * [cx0 - 1, cx1].  The union is staged in LDS in tiles of TILE records,
 * dealt to the tiles round robin in groups of TGRP = 8 (tile_union_index); each
 * lane walks its own neighbour list tile by tile.  Pair terms never
 * scatter: every pass is owner-computes.
 *
 * Neighbour lists (built by k_build_t, read by every pass): rows of four
 * 16-bit entries k*TILE + t (slot t of tile k), row r of the target at
 * position p of vorder[] stored at lu4[r*npad + p].  Within a row the
 * entries ascend and lie in one tile; unused entries hold NO_ENTRY.  The 64 targets of a wave hold the
 * same number of rows for every tile (rows padded to the wave's maximum), so
 * a wave walks its lists in lock-step: every row load is one coalesced
 * 512-byte read.  Distances are not stored: the reference's frozen distance
 * (sqrt of the d^2 of the last build, src/Resid.cpp:288-289) is recomputed
 * from the positions of the last build (Aux::xb).
 *
 * Reference correspondence (file:line):
 *   k_build_t      update_neighbours            src/Neighbours.cpp:7-47
 *   k_dsph_t       dSPH_PreStep                 src/Shifting.cpp:12-123
 *   k_surface1_t   Detect_Surface loop 1        src/Geometry.cpp:26-143
 *                  + dissipation_terms          src/Shifting.cpp:126-186
 *   k_surface2_t   Detect_Surface loops 2, 3    src/Geometry.cpp:145-277
 *   k_shift_t      particle_shift               src/Shifting.cpp:189-290
 *   k_bnd_*        boundary solvers             src/Resid.cpp:21-186, src/Newmark_Beta.cpp:69-132
 *   k_force_t      get_acc_and_Rrho_on_i        src/Resid.cpp:243-418
 *   (k_force_m: the same in FP32 pair terms; the k_*_m kernels are the
 *   mixed-precision forms of the k_*_t ones)
 *   k_nb_update    Do_NB_Iter update            src/Newmark_Beta.cpp:137-240
 */
#include "device_common.hpp"

namespace fj {

constexpr int MAXSEG = 9;

/* Union of block b (see file header). seg_a[k] is the order[] start of
 * segment k, pre[k] its first union index; the target positions in order[]
 * are [p0, p1). */
template <int D>
__device__ void union_of_block(const TileGrid& tg, int64_t b, int32_t* seg_a, int32_t* pre, int& nseg, int& p0,
                               int& p1)
{
    const int64_t row = b / tg.nchunk;
    const int q = (int)(b - row * tg.nchunk);
    const int cx0 = q * tg.C;
    const int cx1 = min(cx0 + tg.C, tg.dims[0]);
    const int64_t rbase = row * tg.dims[0];
    p0 = tg.start[rbase + cx0];
    p1 = tg.start[rbase + cx1];
    const int cy = (int)(row % tg.dims[1]);
    const int cz = (D == 3) ? (int)(row / tg.dims[1]) : 0;
    const int xa = max(cx0 - 1, 0), xb = min(cx1, tg.dims[0] - 1);
    nseg = 0;
    pre[0] = 0;
    const int zlo = (D == 3) ? max(cz - 1, 0) : 0, zhi = (D == 3) ? min(cz + 1, tg.dims[2] - 1) : 0;
    for (int z = zlo; z <= zhi; ++z)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, tg.dims[1] - 1); ++y) {
            const int64_t r2 = ((int64_t)z * tg.dims[1] + y) * tg.dims[0];
            const int32_t a = tg.start[r2 + xa], e = tg.start[r2 + xb + 1];
            seg_a[nseg] = a;
            pre[nseg + 1] = pre[nseg] + (e - a);
            ++nseg;
        }
}

__device__ __forceinline__ int32_t union_slot(const TileGrid& tg, const int32_t* seg_a, const int32_t* pre, int nseg,
                                              int u)
{
    int k = 0;
    while (k + 1 < nseg && u >= pre[k + 1]) ++k;
    return tg.order[seg_a[k] + (u - pre[k])];
}

/* Union indices are dealt to the LDS tiles in groups of TGRP consecutive
 * indices, round robin: tile k holds groups k, k + NT, k + 2 NT, ... of the
 * NT = ceil(U / TILE) tiles.  Every tile then samples the whole union
 * evenly, so each target's neighbours spread evenly over the tiles and the
 * lock-step rows keep most lanes of a wave busy in every tile (with tiles
 * of consecutive indices a tile covers one slab of the union, and a wave's
 * per-tile list lengths differ several-fold between targets near and far
 * from it).  Tile-local slot t of tile k is union index
 * ((t / TGRP) NT + k) TGRP + t % TGRP; a list entry stores k TILE + t, so
 * entries ascend with the tile. */
#ifndef FJ_TGRP
#define FJ_TGRP 8
#endif
constexpr int TGRP = FJ_TGRP;
static_assert(TILE % TGRP == 0, "tile slots come in whole groups");

__device__ __forceinline__ int n_tiles(int U) { return (U + TILE - 1) / TILE; }

__device__ __forceinline__ int tile_union_index(int k, int t, int NT) { return ((t / TGRP) * NT + k) * TGRP + t % TGRP; }

/* slots of tile k (the last group of the union may be partial: its slots
 * past U are padding) */
__device__ __forceinline__ int tile_len(int k, int U, int NT)
{
    const int ng = (U + TGRP - 1) / TGRP;
    return k < ng ? ((ng - 1 - k) / NT + 1) * TGRP : 0;
}

/* union index of a list entry */
__device__ __forceinline__ int entry_union_index(uint32_t e, int NT)
{
    return tile_union_index(int(e / TILE), int(e % TILE), NT);
}

/* Block id of the block that owns slot i at the last build. */
__device__ __forceinline__ int64_t block_of_slot(const TileGrid& tg, int64_t i)
{
    const int32_t lin = tg.cell[i];
    const int64_t row = lin / tg.dims[0];
    const int cx = lin - (int)(row * tg.dims[0]);
    return row * tg.nchunk + cx / tg.C;
}

/* Group of block b (independent simulations): the groups occupy disjoint
 * cell layers along the last axis, and a block lies in one cell row. */
template <int D>
__device__ __forceinline__ int32_t block_group(const TileGrid& tg, int64_t b)
{
    const int64_t row = b / tg.nchunk;
    const int64_t layer = (D == 3) ? row / tg.dims[1] : row;
    return tg.lgrp[layer];
}

/* Shared block prologue: union table in LDS, uniform p0/p1/U, and the
 * block b the workgroup runs. */
struct BlockHead {
    int32_t seg_a[MAXSEG];
    int32_t pre[MAXSEG + 1];
    int32_t nseg, p0, p1;
    int64_t b;
};

template <int D>
__device__ __forceinline__ void block_head(const TileGrid& tg, BlockHead& h)
{
    if (threadIdx.x == 0) {
        int nseg, p0, p1;
        /* workgroup w runs block w: dispatch order is cell order, so the
         * blocks in flight cover a band of neighbouring cell rows */
        const int64_t b = blockIdx.x;
        h.b = b;
        union_of_block<D>(tg, b, h.seg_a, h.pre, nseg, p0, p1);
        h.nseg = nseg;
        h.p0 = p0;
        h.p1 = p1;
    }
    __syncthreads();
}

/* The slots a thread stages into LDS tile tk (t = threadIdx.x and
 * threadIdx.x + TPB; -1 where there is none).  Loaded one tile ahead, so
 * the order[] lookup of the next tile is in flight while the current
 * tile's pairs run and staging after the barrier waits only for the data. */
static_assert(TILE <= 2 * TPB, "a thread stages at most two slots per tile");
struct TileSlots {
    int32_t sl[2];
    __device__ __forceinline__ void load(const TileGrid& tg, const BlockHead& h, int tk, int NT, int U)
    {
        const int nt = tk < NT ? tile_len(tk, U, NT) : 0;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int t = threadIdx.x + r * TPB;
            sl[r] = -1;
            if (t < nt) {
                const int u = tile_union_index(tk, t, NT);
                if (u < U) sl[r] = union_slot(tg, h.seg_a, h.pre, h.nseg, u);
            }
        }
    }
    __device__ __forceinline__ int32_t at(int t) const { return t < TPB ? sl[0] : sl[1]; }
};

/* x - ref as an FP32 value hi and its FP32 rounding residual lo (hi + lo
 * = x - ref to about 1e-14 relative).  A pair difference is formed as
 * (xj_hi - hi) - lo: the target's own rounding then cancels, and only the
 * neighbour's (independent from pair to pair) remains, instead of an
 * error common to every pair of the target that sums coherently. */
__device__ __forceinline__ void split_rel(double x, double ref, float& hi, float& lo)
{
    const double d = x - ref;
    hi = float(d);
    lo = float(d - double(hi));
}

/* ------------------------------------------------------------ list rows */

/* Lock-step walk of one target's rows: `row` advances in every lane of the
 * wave together; the next row is in flight while the current one is used.
 * Rows past nrows read as all-NO_ENTRY.  Every lane with a list walks it,
 * also a lane whose target the pass does not compute (`on` false): the
 * build pads each tile's rows to the wave's largest count, and the wave
 * leaves a tile only when no lane has an entry of it left, so a lane that
 * dropped out could leave the others on padding rows of the tile. */
struct Rows {
    const uint64_t* rp; /* the next row (row nxt of the target) */
    int64_t stride;
    int nxt, nrows;
    bool on;
    uint64_t u, un;
    __device__ __forceinline__ void init(const NbrList& nl, int64_t p, int nrows_, bool on_)
    {
        on = on_;
        nrows = nrows_;
        stride = nl.npad;
        rp = nl.lu4 + p;
        u = nrows > 0 ? *rp : ~0ull;
        rp += stride;
        nxt = 1;
        un = nrows > 1 ? *rp : ~0ull;
    }
    __device__ __forceinline__ void advance(const NbrList&)
    {
        u = un;
        rp += stride;
        ++nxt;
        un = nxt < nrows ? *rp : ~0ull;
    }
};

/* FJ_TILE_ENTRIES_BEGIN(cu, nl, base) { body } FJ_TILE_ENTRIES_END(cu, nl)
 * runs body for every list entry of the lane's target in the union tile
 * [base, base + TILE), with `t` the tile-local index.  A row belongs to the
 * tile when its first index does; padding rows read NO_ENTRY first, so a
 * lane is idle in them, as is a lane with `on` false.  The wave leaves the loop at the first row where
 * no lane has an entry of the tile, which is the first row of the next
 * tile in every lane.  The 4 entries of a row go through one copy of the
 * body (a rolled loop); the body is written inline, not as a closure, so
 * per-lane accumulator arrays stay in registers. */
#define FJ_TILE_ENTRIES_BEGIN(cu, nl, base)                                               \
    for (;;) {                                                                            \
        const bool fj_act = (uint32_t((cu).u) & 0xFFFFu) < uint32_t((base) + TILE);        \
        if (!__any(fj_act)) break;                                                        \
        if (fj_act && (cu).on) {                                                          \
            uint64_t fj_u = (cu).u;                                                       \
            _Pragma("unroll 1") for (int fj_q = 0; fj_q < 4; ++fj_q)                     \
            {                                                                             \
                const uint32_t fj_uq = uint32_t(fj_u) & 0xFFFFu;                          \
                fj_u >>= 16;                                                              \
                if (fj_uq == NO_ENTRY) continue;                                          \
                const int t = int(fj_uq) - (base);

#define FJ_TILE_ENTRIES_END(cu, nl) \
            }                       \
        }                           \
        (cu).advance(nl);           \
    }

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_minf(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float wave_maxf(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}


/* ------------------------------------------------------------ list build */

/* A candidate is a neighbour when its FP64 squared distance from the
 * target is below sr (src/Neighbours.cpp:44, nanoflann's strict test).
 * The test runs first in FP32 on coordinates relative to the block's first
 * target, each rounded once to FP32: with every coordinate within L of that
 * point, the FP32 squared distance near sr is off by at most about
 * 2^-23 (L / H) sr plus a few ulp, and L < (C + 2) cs sqrt(3) < 1000 H for
 * any tile grid (C <= 256 cells of size 2H), so below (1 - 1e-3) sr it is
 * surely in, above (1 + 1e-3) sr surely out, and only the thin shell
 * between is decided by the exact FP64 test.  The lists are those of the
 * FP64 test.  The candidate loop only marks what is not surely out; the
 * in/shell split and the FP64 test run on the marked ones (the ~240
 * neighbours of ~2000 candidates), in ascending order. */
template <int D>
__global__ __launch_bounds__(TPB) void k_build_t(const double* __restrict__ x, TileGrid tg, double sr, NbrList nl,
                                                 DevScalars* sc)
{
    __shared__ double tp[D][TILE];
    __shared__ float4 tf[(TILE + 31) / 32 * 32]; /* read in chunks of 32 */
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    if (h.p0 == h.p1) return;
    if (tg.gbuild) {
        /* a partial build (independent simulations): other groups keep
         * their lists */
        const int32_t g = block_group<D>(tg, h.b);
        if (g < 0 || !tg.gbuild[g]) return;
    }
    if (NT * TILE > int(NO_ENTRY)) {
        if (threadIdx.x == 0) atomicMax(&sc->overflow, U);
        return;
    }
    double xref[D];
    {
        const int64_t i0 = tg.order[h.p0];
#pragma unroll
        for (int a = 0; a < D; ++a) xref[a] = x[i0 * D + a];
    }
    const float sr_lo = float(sr * (1.0 - 1e-3)), sr_hi = float(sr * (1.0 + 1e-3));
    const int rcap = nl.cap >> 2;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        if (active) {
            const int32_t sl = tg.order[p];
            tg.vorder[p] = sl;
            tg.pos[sl] = p;
        }
        const int64_t i = active ? tg.order[p] : 0;
        double xi[D];
        float xit[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = active ? x[i * D + a] : xref[a];
            xit[a] = float(xi[a] - xref[a]);
        }
        /* bounding box of the wave's targets (FP32, relative to xref): a
         * candidate farther than the cut-off from it is no neighbour of any
         * of them */
        float blo[3] = {0.f, 0.f, 0.f}, bhi[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < D; ++a) {
            blo[a] = wave_minf(active ? xit[a] : HUGE_VALF);
            bhi[a] = wave_maxf(active ? xit[a] : -HUGE_VALF);
        }
        /* the target's own slot in the union (tile k_self, slot t_self):
         * the segment of the block's own cell row holds its order position */
        int k_self = -1, t_self = 0;
        if (active) {
            int sg = 0;
            while (sg + 1 < h.nseg && !(p >= h.seg_a[sg] && p < h.seg_a[sg] + (h.pre[sg + 1] - h.pre[sg]))) ++sg;
            const int u = h.pre[sg] + (p - h.seg_a[sg]);
            const int g = u / TGRP;
            k_self = g % NT;
            t_self = (g / NT) * TGRP + u % TGRP;
        }
        int nrow = 0, nreal = 0;
        uint64_t* rowp = nl.lu4 + p; /* row nrow of the target */
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int u = tile_union_index(tk, t, NT);
                if (u >= U) {
                    /* padding: far away, never a neighbour */
                    tf[t] = make_float4(1e30f, 1e30f, 1e30f, -2.f);
                    continue;
                }
                const int32_t sl = union_slot(tg, h.seg_a, h.pre, h.nseg, u);
                float pa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    tp[a][t] = x[(int64_t)sl * D + a];
                    pa[a] = float(tp[a][t] - xref[a]);
                }
                tf[t] = make_float4(pa[0], pa[1], pa[2], 0.f);
            }
            __syncthreads();
            /* groups of 4 slots with a candidate in reach of the wave's
             * box (squared distance to it at most sr_hi, which every
             * neighbour's FP32 squared distance to its target meets): one
             * bit per group, wave-uniform (tiles hold at most 128 groups) */
            uint64_t gmask[2];
#pragma unroll
            for (int hgrp = 0; hgrp < 2; ++hgrp) {
                const int g = hgrp * 64 + (threadIdx.x & 63);
                bool reach = false;
                if (4 * g < nt) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float4 P = tf[4 * g + e];
                        const float pp[3] = {P.x, P.y, P.z};
                        float d2b = 0.f;
#pragma unroll
                        for (int a = 0; a < D; ++a) {
                            const float da = fmaxf(fmaxf(blo[a] - pp[a], pp[a] - bhi[a]), 0.f);
                            d2b += da * da;
                        }
                        reach = reach || d2b <= sr_hi;
                    }
                }
                gmask[hgrp] = __ballot(reach);
            }
            int rows_t = 0;
            if (active) {
                int q = 0;
                uint64_t cu = ~0ull;
                /* 32 candidates at a time: a branch-free FP32 test into two
                 * bit masks, those not surely out (near) and those surely
                 * in; the FP64 test decides the rest of near (the shell),
                 * and the target's own slot is masked out; the neighbours
                 * are then emitted in ascending order */
                for (int c0 = 0; c0 < nt; c0 += 32) {
                    const uint32_t gm = uint32_t(gmask[c0 >> 8] >> ((c0 >> 2) & 63)) & 0xFFu;
                    if (gm == 0u) continue;
                    uint32_t near = 0u, in = 0u;
                    /* candidates in descending order, each shifted in at
                     * bit 0, so candidate k ends at bit k */
#pragma unroll
                    for (int gj = 7; gj >= 0; --gj) {
                        if (!((gm >> gj) & 1u)) {
                            near <<= 4;
                            in <<= 4;
                            continue;
                        }
#pragma unroll
                        for (int e = 3; e >= 0; --e) {
                            const float4 P = tf[c0 + 4 * gj + e];
                            const float d0 = P.x - xit[0], d1 = P.y - xit[1], d2f = P.z - xit[2];
                            const float rr = d0 * d0 + d1 * d1 + d2f * d2f;
                            near = near + near + uint32_t(rr <= sr_hi);
                            in = in + in + uint32_t(rr < sr_lo);
                        }
                    }
                    const int rem = nt - c0;
                    near &= rem >= 32 ? ~0u : ((1u << rem) - 1u);
                    in &= near;
                    if (tk == k_self && unsigned(t_self - c0) < 32u) {
                        const uint32_t sb = 1u << (t_self - c0);
                        near &= ~sb;
                        in &= ~sb;
                    }
                    for (uint32_t shell = near & ~in; shell;) {
                        const int k = __builtin_ctz(shell);
                        shell &= shell - 1u;
                        double xj[D];
#pragma unroll
                        for (int a = 0; a < D; ++a) xj[a] = tp[a][c0 + k];
                        if (exact_dist2<D>(xi, xj) < sr) in |= 1u << k;
                    }
                    nreal += __builtin_popcount(in);
                    while (in) {
                        const int k = __builtin_ctz(in);
                        in &= in - 1u;
                        cu = (cu >> 16) | (uint64_t(base + c0 + k) << 48);
                        if (++q == 4) {
                            if (nrow < rcap) *rowp = cu;
                            ++nrow;
                            rowp += nl.npad;
                            ++rows_t;
                            q = 0;
                            cu = ~0ull;
                        }
                    }
                }
                if (q > 0) {
                    /* close the tile's last row: entries to the low slots */
                    cu = (cu >> (16 * (4 - q))) | (~0ull << (16 * q));
                    if (nrow < rcap) *rowp = cu;
                    ++nrow;
                    rowp += nl.npad;
                    ++rows_t;
                }
            }
            /* pad every lane of the wave to the wave's row count of this tile */
            const int wrows = wave_max(rows_t);
            if (active) {
                for (int k = rows_t; k < wrows; ++k) {
                    if (nrow < rcap) *rowp = ~0ull;
                    ++nrow;
                    rowp += nl.npad;
                }
            }
        }
        if (active) {
            nl.count[i] = nrow;
            nl.nnb[i] = nreal;
            atomicMax(&sc->max_count, nrow * 4);
        }
    }
}

void launch_build_lists(int dim, const double* x, int64_t n, TileGrid tg, double sr, NbrList nl, DevScalars* sc,
                        hipStream_t st)
{
    (void)n;
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_build_t<2>, g, b, 0, st, x, tg, sr, nl, sc);
    else hipLaunchKernelGGL(k_build_t<3>, g, b, 0, st, x, tg, sr, nl, sc);
}

/* f(j) for every neighbour slot j of slot i, one lane per slot (the passes
 * over the thin boundary layer and the host dump walk lists this way). */
template <int D, typename F>
__device__ __forceinline__ void walk_lane(const TileGrid& tg, const NbrList& nl, int64_t i, F&& f)
{
    int32_t seg_a[MAXSEG], pre[MAXSEG + 1];
    int nseg, p0, p1;
    union_of_block<D>(tg, block_of_slot(tg, i), seg_a, pre, nseg, p0, p1);
    const int NT = n_tiles(pre[nseg]);
    const int64_t p = tg.pos[i];
    const int nrows = min(nl.count[i], nl.cap >> 2);
    for (int r = 0; r < nrows; ++r) {
        const uint64_t u = nl.lu4[(int64_t)r * nl.npad + p];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t uq = uint32_t(u >> (16 * q)) & 0xFFFFu;
            if (uq != NO_ENTRY) f(union_slot(tg, seg_a, pre, nseg, entry_union_index(uq, NT)));
        }
    }
}

/* Global slots of every slot's neighbours, compacted: entry k of slot i at
 * out[k*npad + i] for k < nnb[i] (for the host dump of the lists). */
template <int D>
__global__ __launch_bounds__(TPB) void k_export(TileGrid tg, NbrList nl, int64_t n, int32_t* out)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    int k = 0;
    walk_lane<D>(tg, nl, i, [&](int32_t j) {
        out[(int64_t)k * nl.npad + i] = j;
        ++k;
    });
}

void launch_export_lists(int dim, TileGrid tg, NbrList nl, int64_t n, int32_t* out, hipStream_t st)
{
    if (dim == 2) hipLaunchKernelGGL(k_export<2>, dim3(grid_of(n)), dim3(TPB), 0, st, tg, nl, n, out);
    else hipLaunchKernelGGL(k_export<3>, dim3(grid_of(n)), dim3(TPB), 0, st, tg, nl, n, out);
}

/* Distance of a pair as the reference uses it: sqrt of the d^2 of the last
 * neighbour build.  FR = false: the positions are still those of the build
 * (xi, xj); FR = true: the build positions of both (xbi, and xb of slot j). */
template <int D, bool FR>
__device__ __forceinline__ void frozen_r(const double* xi, const double* xj, const double* xbi,
                                         const double* __restrict__ xb, int32_t j, double& d2, double& r)
{
    if constexpr (FR) {
        double xbj[D];
#pragma unroll
        for (int a = 0; a < D; ++a) xbj[a] = xb[(int64_t)j * D + a];
        d2 = exact_dist2<D>(xbi, xbj);
    } else {
        d2 = exact_dist2<D>(xi, xj);
    }
    r = sqrt(d2);
}

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

/* One pair of dSPH_PreStep in FP64 (src/Shifting.cpp:30-66): adds to the
 * correction matrices (upper triangles), the density-gradient and normal
 * sums, the kernel sum and colour; returns the pair's share of npd. */
template <int D>
__device__ __forceinline__ double dsph_pair(const Consts& c, double inv_H, const double* xi, const double* xj, double r,
                                            double volj, double rhoj, bool nb, double rhoi,
                                            double (&Lm)[D * (D + 1) / 2], double (&Lnb)[D * (D + 1) / 2], double* gR,
                                            double* nr, double& kernsum, double& colour)
{
    double Rji[D];
#pragma unroll
    for (int a = 0; a < D; ++a) Rji[a] = xj[a] - xi[a];
    double W, fg;
    wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
    double G[D];
#pragma unroll
    for (int a = 0; a < D; ++a) G[a] = -fg * Rji[a];
    const double drho = volj * (rhoj - rhoi);
    int k = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double vr = volj * Rji[a];
#pragma unroll
        for (int b2 = a; b2 < D; ++b2, ++k) {
            const double tt = vr * G[b2];
            Lm[k] -= tt;
            if (nb) Lnb[k] -= tt;
        }
        gR[a] += drho * G[a];
    }
    if (!nb) return 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) nr[a] += volj * G[a];
    kernsum += W;
    colour += volj * W;
    return W;
}

/* Per-slot outputs of the dSPH_PreStep pair sums. */
template <int D>
__device__ __forceinline__ void dsph_out(DevState& s, Aux& ax, int64_t i, const double (&Lm)[D * (D + 1) / 2],
                                         const double (&Lnb)[D * (D + 1) / 2], const double* gR, const double* nr,
                                         double kernsum, double colour)
{
    int k = 0;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b2 = a; b2 < D; ++b2, ++k) {
            s.L[i * D * D + a * D + b2] = Lm[k];
            s.L[i * D * D + b2 * D + a] = Lm[k];
            ax.lnb[i * D * D + a * D + b2] = Lnb[k];
            ax.lnb[i * D * D + b2 * D + a] = Lnb[k];
        }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        s.gradRho[i * D + a] = gR[a];
        s.norm[i * D + a] = nr[a];
    }
    s.kernsum[i] = kernsum;
    s.colour[i] = colour;
}

/* Pair sums of dSPH_PreStep (src/Shifting.cpp:12-123): the correction
 * matrix L over all neighbours and over non-boundary neighbours (both
 * symmetric, sum_j V_j W'(r)/r R_ji R_ji^T, accumulated on the upper
 * triangle), the density-gradient and normal sums, the kernel sum and the
 * colour function, and the per-block partial of npd.  k_dsph_fin turns
 * them into L^-1, the eigenvalues and the corrected gradients. */
template <int D, bool FR>
__global__ __launch_bounds__(TPB) void k_dsph_t(Consts c, DevState s, Aux ax, TileGrid tg, NbrList nl, double* partial)
{
    constexpr int NS = D * (D + 1) / 2;
    __shared__ double tp[D][TILE];
    __shared__ double tvol[TILE], trho[TILE];
    __shared__ int32_t tnb[TILE], tsl[FR ? TILE : 1];
    __shared__ BlockHead h;
    __shared__ double lds[NWAVE];
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double npd_acc = 0.0;
    const double inv_H = 1.0 / c.H;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        double xi[D], xbi[D], Lm[NS], Lnb[NS], gR[D], nr[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = active ? s.x[i * D + a] : 0.0;
            xbi[a] = (FR && active) ? ax.xb[i * D + a] : 0.0;
            gR[a] = 0.0;
            nr[a] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) { Lm[k] = 0.0; Lnb[k] = 0.0; }
        const double rhoi = active ? s.rho[i] : 0.0;
        double kernsum = c.Wc, colour = 0.0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, active);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
#pragma unroll
                for (int a = 0; a < D; ++a) tp[a][t] = s.x[(int64_t)sl * D + a];
                tvol[t] = ax.vol[sl];
                trho[t] = s.rho[sl];
                tnb[t] = s.b[sl] > FJSPH_PISTON;
                if constexpr (FR) tsl[t] = sl;
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                double xj[D], d2, r;
#pragma unroll
                for (int a = 0; a < D; ++a) xj[a] = tp[a][t];
                frozen_r<D, FR>(xi, xj, xbi, ax.xb, FR ? tsl[FR ? t : 0] : 0, d2, r);
                npd_acc += dsph_pair<D>(c, inv_H, xi, xj, r, tvol[t], trho[t], tnb[t] != 0, rhoi, Lm, Lnb, gR, nr,
                                        kernsum, colour);
            FJ_TILE_ENTRIES_END(cu, nl)
        }
        if (active) dsph_out<D>(s, ax, i, Lm, Lnb, gR, nr, kernsum, colour);
    }
    const double bs = block_reduce(npd_acc, OpSum(), lds);
    if (threadIdx.x == 0) partial[h.b] = bs;
}

/* Mixed-precision form of k_dsph_t: FP32 pair terms on coordinates relative
 * to the block's first target, summed in FP32 over one LDS tile and added
 * to FP64 totals after each tile (about 40 pairs per tile and target);
 * 20 B of LDS per candidate (36 B with build positions).  A target with
 * fewer than 2 D neighbours takes its pairs in FP64 from the global state
 * (dsph_pair), so the rank test of its correction matrix, which decides
 * between L^-1 and the identity, sees the matrix of the FP64 path. */
#ifndef FJ_DSPH_OCC
#define FJ_DSPH_OCC __attribute__((amdgpu_waves_per_eu(5, 8)))
#endif
template <int D, bool FR>
__global__ __launch_bounds__(TPB) FJ_DSPH_OCC void k_dsph_m(Consts c, DevState s, Aux ax, TileGrid tg, NbrList nl, double* partial)
{
    constexpr int NS = D * (D + 1) / 2;
    __shared__ float4 tA[TILE];            /* x_j - xref, +V_j (-V_j: j is PISTON or BOUND) */
    __shared__ float tR[TILE];             /* rho_j - rref */
    __shared__ float4 tX[FR ? TILE : 1];   /* xb_j - xbref */
    __shared__ BlockHead h;
    __shared__ double lds[NWAVE];
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double xref[D], xbref[D], rref;
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
        rref = s.rho[i0];
    }
    const float Ht = float(c.H), inv_Hf = float(1.0 / c.H), Wcf = float(c.Wc);
    const double inv_H = 1.0 / c.H;
    double npd_acc = 0.0;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool slow = active && nl.nnb[i] < 2 * D;
        float xit[D], xlo[D], xbit[D], xblo[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            split_rel(active ? s.x[i * D + a] : xref[a], xref[a], xit[a], xlo[a]);
            split_rel((FR && active) ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
        }
        const double rhoi = active ? s.rho[i] : rref;
        float drhoi, rlo;
        split_rel(rhoi, rref, drhoi, rlo);
        double Lm[NS], Lnb[NS], gR[D], nr[D], kernsum = c.Wc, colour = 0.0;
        float fLm[NS], fLnb[NS], fgR[D], fnr[D], fks = 0.f, fcol = 0.f;
#pragma unroll
        for (int k = 0; k < NS; ++k) { Lm[k] = 0.0; Lnb[k] = 0.0; fLm[k] = 0.f; fLnb[k] = 0.f; }
#pragma unroll
        for (int a = 0; a < D; ++a) { gR[a] = 0.0; nr[a] = 0.0; fgR[a] = 0.f; fnr[a] = 0.f; }
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, active && !slow);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                const float4 r2 = ax.fs2[sl]; /* V, .., m negative for PISTON and BOUND */
                tA[t] = make_float4(pa[0], pa[1], pa[2], copysignf(r2.x, r2.z));
                tR[t] = float(s.rho[sl] - rref);
                if constexpr (FR) tX[t] = make_float4(xa[0], xa[1], xa[2], 0.f);
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = tA[t];
                const float pj[3] = {A.x, A.y, A.z};
                float Rji[D], rr;
#pragma unroll
                for (int a = 0; a < D; ++a) Rji[a] = (pj[a] - xit[a]) - xlo[a];
                if constexpr (FR) {
                    const float4 X = tX[FR ? t : 0];
                    const float xbj[3] = {X.x, X.y, X.z};
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rji[0] * Rji[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rji[a] * Rji[a];
                }
                float W, fg;
                wendland<float>(fast_sqrt<float>(rr), Ht, inv_Hf, Wcf, W, fg);
                const bool nb = A.w > 0.f;
                const float volj = fabsf(A.w);
                const float nbf = nb ? 1.f : 0.f;
                float G[D];
#pragma unroll
                for (int a = 0; a < D; ++a) G[a] = -fg * Rji[a];
                const float drho = volj * ((tR[t] - drhoi) - rlo);
                int k = 0;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const float vr = volj * Rji[a];
#pragma unroll
                    for (int b2 = a; b2 < D; ++b2, ++k) {
                        const float tt = vr * G[b2];
                        fLm[k] -= tt;
                        fLnb[k] -= nbf * tt;
                    }
                    fgR[a] += drho * G[a];
                    fnr[a] += nbf * volj * G[a];
                }
                fks += nbf * W;
                fcol += nbf * volj * W;
            FJ_TILE_ENTRIES_END(cu, nl)
            /* the tile's FP32 partial sums into the FP64 totals */
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                Lm[k] += double(fLm[k]);
                Lnb[k] += double(fLnb[k]);
                fLm[k] = 0.f;
                fLnb[k] = 0.f;
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                gR[a] += double(fgR[a]);
                nr[a] += double(fnr[a]);
                fgR[a] = 0.f;
                fnr[a] = 0.f;
            }
            kernsum += double(fks);
            colour += double(fcol);
            npd_acc += double(fks);
            fks = 0.f;
            fcol = 0.f;
        }
        if (slow) {
            double xi[D], xbi[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                xi[a] = s.x[i * D + a];
                xbi[a] = FR ? ax.xb[i * D + a] : 0.0;
            }
            const int nrows = min(nl.count[i], nl.cap >> 2);
            for (int r = 0; r < nrows; ++r) {
                const uint64_t e4 = nl.lu4[(int64_t)r * nl.npad + p];
                for (int q = 0; q < 4; ++q) {
                    const uint32_t e = uint32_t(e4 >> (16 * q)) & 0xFFFFu;
                    if (e == NO_ENTRY) continue;
                    const int64_t j = union_slot(tg, h.seg_a, h.pre, h.nseg, entry_union_index(e, NT));
                    double xj[D], d2, rj;
#pragma unroll
                    for (int a = 0; a < D; ++a) xj[a] = s.x[j * D + a];
                    frozen_r<D, FR>(xi, xj, xbi, ax.xb, (int32_t)j, d2, rj);
                    npd_acc += dsph_pair<D>(c, inv_H, xi, xj, rj, ax.vol[j], s.rho[j], s.b[j] > FJSPH_PISTON, rhoi,
                                            Lm, Lnb, gR, nr, kernsum, colour);
                }
            }
        }
        if (active) dsph_out<D>(s, ax, i, Lm, Lnb, gR, nr, kernsum, colour);
    }
    const double bs = block_reduce(npd_acc, OpSum(), lds);
    if (threadIdx.x == 0) partial[h.b] = bs;
}

/* Per-slot tail of dSPH_PreStep: L^-1 (column-pivoting QR, identity when
 * L is rank deficient), the minimum eigenvalues, the corrected density
 * gradient and normal, the colour gradient factor, and get_aero_velocity
 * for the constant-velocity source. */
template <int D>
__global__ __launch_bounds__(TPB) void k_dsph_fin(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                  const int32_t* __restrict__ nnb, int64_t n, int with_aero,
                                                  int mixed)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    double Lm[D][D], Lw[D][D], Linv[D][D], tmp[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b2 = 0; b2 < D; ++b2) {
            Lm[a][b2] = s.L[i * D * D + a * D + b2];
            Lw[a][b2] = Lm[a][b2];
            Linv[a][b2] = (a == b2) ? 1.0 : 0.0;
        }
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
        double Lnb[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) Lnb[a][b2] = ax.lnb[i * D * D + a * D + b2];
        lam = min_eig(Lm);
        lam_nb = min_eig(Lnb);
    }
    s.lam[i] = lam;
    s.lam_nb[i] = lam_nb;
    double gR[D], nr[D];
    load_vec<D>(s.gradRho, i, gR);
    load_vec<D>(s.norm, i, nr);
    float gf[3] = {0.f, 0.f, 0.f};
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
        gf[a] = float(g);
    }
    /* the FP32 staging record of the surface and dissipation pass */
    if (mixed) ax.fs1[i] = make_float4(gf[0], gf[1], gf[2], float(lam));
    const double colour = s.colour[i];
    s.colourG[i] = lam > 0.7 ? 2.0 : 2.0 * fmax(1.0, 1.0 / (2.0 * colour));
    if (with_aero && c.aero_source == FJSPH_CONST_VEL && part_id[i] >= c.n_bound)
        aero_velocity_one<D>(c, s, i, lam_nb, nnb[i] + 1);
}

void launch_dsph(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                 int64_t n, int r_from_x, int with_aero_vel, double* partial, DevScalars* sc, hipStream_t st)
{
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if (nl.f32) {
        if (dim == 2) {
            if (r_from_x) hipLaunchKernelGGL((k_dsph_m<2, false>), g, b, 0, st, c, s, a, tg, nl, partial);
            else hipLaunchKernelGGL((k_dsph_m<2, true>), g, b, 0, st, c, s, a, tg, nl, partial);
        } else {
            if (r_from_x) hipLaunchKernelGGL((k_dsph_m<3, false>), g, b, 0, st, c, s, a, tg, nl, partial);
            else hipLaunchKernelGGL((k_dsph_m<3, true>), g, b, 0, st, c, s, a, tg, nl, partial);
        }
    } else if (dim == 2) {
        if (r_from_x) hipLaunchKernelGGL((k_dsph_t<2, false>), g, b, 0, st, c, s, a, tg, nl, partial);
        else hipLaunchKernelGGL((k_dsph_t<2, true>), g, b, 0, st, c, s, a, tg, nl, partial);
    } else {
        if (r_from_x) hipLaunchKernelGGL((k_dsph_t<3, false>), g, b, 0, st, c, s, a, tg, nl, partial);
        else hipLaunchKernelGGL((k_dsph_t<3, true>), g, b, 0, st, c, s, a, tg, nl, partial);
    }
    launch_sum_partials_scaled(partial, (int)tg.nblocks, 1, 1.0 / double(n), &sc->npd, st);
    if (dim == 2)
        hipLaunchKernelGGL(k_dsph_fin<2>, dim3(grid_of(n)), b, 0, st, c, s, a, part_id, nl.nnb, n, with_aero_vel,
                           nl.f32);
    else
        hipLaunchKernelGGL(k_dsph_fin<3>, dim3(grid_of(n)), b, 0, st, c, s, a, part_id, nl.nnb, n, with_aero_vel,
                           nl.f32);
}

template <int D>
__global__ __launch_bounds__(TPB) void k_aero_vel(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                  NbrList nl, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    aero_velocity_one<D>(c, s, i, s.lam_nb[i], nl.nnb[i] + 1);
}

void launch_aero_velocity(int dim, const Consts& c, DevState s, const int32_t* part_id, NbrList nl, int64_t n,
                          hipStream_t st)
{
    if (c.aero_source != FJSPH_CONST_VEL) return;
    if (dim == 2) hipLaunchKernelGGL(k_aero_vel<2>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, part_id, nl, n);
    else hipLaunchKernelGGL(k_aero_vel<3>, dim3(grid_of(n)), dim3(TPB), 0, st, c, s, part_id, nl, n);
}

/* -------------------------------------- Detect_Surface + dissipation_terms */

/* Marrone et al. (2010) point-T test of one pair, src/Geometry.cpp:43-98:
 * does neighbour j (at xj, frozen distance r) fall in the test region of
 * point T = xi + h133 nhat? */
template <int D>
__device__ __forceinline__ bool marrone_pair(const double* xi, const double* xj, double r, const double* nhat,
                                             const double* tau_hat, const double* pT, double h133, double r_lim)
{
    double xjT[D];
#pragma unroll
    for (int a = 0; a < D; ++a) xjT[a] = xj[a] - pT[a];
    if (r >= r_lim) return sqrt(dotd<D>(xjT, xjT)) < h133;
    if (D == 2) return fabs(dotd<D>(nhat, xjT)) + fabs(dotd<D>(tau_hat, xjT)) < h133;
    double mr[D];
#pragma unroll
    for (int a = 0; a < D; ++a) mr[a] = -(xi[a] - xj[a]) / r;
    return acos(dotd<D>(nhat, mr)) < kPi / 4.0;
}

/* The FP64 point-T test of pair (i, j) from the global state: normal,
 * test point and frozen distance rebuilt as k_surface1_t forms them. */
/* surface flag of a target whose point-T test awaits the FP64 decision:
 * kSurfUndecided + the number of undecided pairs (kSurfUndecided + 5: more
 * than the four whose slots Aux::spend keeps) */
constexpr int32_t kSurfUndecided = 2;
constexpr int kSurfPendMax = 4;

template <int D, bool FR>
__device__ __noinline__ bool marrone_exact(const Consts& c, const DevState& s, const Aux& ax, int64_t i, int64_t j)
{
    const double h133 = 1.33 * c.particle_step;
    const double r_lim = sqrt(2.0) * h133;
    double xi[D], xj[D], ni[D], nhat[D], tau_hat[D], pT[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        xi[a] = s.x[i * D + a];
        xj[a] = s.x[j * D + a];
        ni[a] = s.norm[i * D + a];
        tau_hat[a] = 0.0;
    }
    normalize<D>(ni, nhat);
    if (D == 2) {
        double tau[D];
        tau[0] = ni[1 % D];
        tau[1 % D] = -ni[0];
        normalize<D>(tau, tau_hat);
    }
#pragma unroll
    for (int a = 0; a < D; ++a) pT[a] = xi[a] + h133 * nhat[a];
    double d2;
    if constexpr (FR) {
        double xbi[D], xbj[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xbi[a] = ax.xb[i * D + a];
            xbj[a] = ax.xb[j * D + a];
        }
        d2 = exact_dist2<D>(xbi, xbj);
    } else {
        d2 = exact_dist2<D>(xi, xj);
    }
    return marrone_pair<D>(xi, xj, sqrt(d2), nhat, tau_hat, pT, h133, r_lim);
}

/* Per-target results of Detect_Surface loop 1 and dissipation_terms
 * (src/Geometry.cpp:100-143, src/Shifting.cpp:175-185). */
template <int D, int MODE>
__device__ __forceinline__ void surface1_out(const Consts& c, DevState& s, int64_t i, bool fluid, int32_t bi,
                                             double lam_nb, double lami, bool found, const double* nrm,
                                             const double* av, double rd, double* norms_tmp, double* nhat)
{
    constexpr bool with_surf = MODE != 2;
    constexpr bool with_diss = MODE != 0;
    if (!fluid) {
        if constexpr (with_surf) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                norms_tmp[i * D + a] = 0.0;
                nhat[i * D + a] = 0.0;
            }
        }
        return;
    }
    if constexpr (with_diss) {
#pragma unroll
        for (int a = 0; a < D; ++a) s.aVisc[i * D + a] = av[a];
        s.deltaD[i] = c.dsph_cont * rd;
    }
    if constexpr (with_surf) {
        int surf;
        if (bi < FJSPH_PIPE) surf = 0;
        else if (lam_nb < 0.2) surf = 1;
        else if (lam_nb < 0.75) surf = found ? 0 : 1;
        else surf = 0;
        s.surf[i] = surf;
        if (bi < FJSPH_PIPE) s.woccl[i] = 1.0;
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
        /* the unit normal of loops 2-3 and particle_shift (the normalised
         * loop-1 normal, as Detect_Surface forms it after its loops) */
        double nh[D];
        normalize<D>(out, nh);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            norms_tmp[i * D + a] = out[a];
            nhat[i * D + a] = nh[a];
        }
    }
}

/* MODE 0: Detect_Surface loop 1 only; 1: with dissipation_terms;
 * 2: dissipation_terms only.  T is the arithmetic of the dissipation terms
 * (FP32 on the mixed path). */
template <int D, typename T, int MODE, bool FR>
__global__ __launch_bounds__(TPB) void k_surface1_t(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                    TileGrid tg, NbrList nl, double* norms_tmp)
{
    constexpr bool with_surf = MODE != 2;
    constexpr bool with_diss = MODE != 0;
    constexpr int TS = with_surf ? TILE : 1, TD = with_diss ? TILE : 1;
    __shared__ double tp[D][TILE], tv[D][TD], tg_[D][TD];
    __shared__ double tvol[TILE], tlam[TS], trho[TD], tir[TD], tm[TD];
    __shared__ int32_t tnb[TD], tsl[FR ? TILE : 1];
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    const double inv_H = 1.0 / c.H;
    const double h133 = 1.33 * c.particle_step;
    const double r_lim = sqrt(2.0) * h133;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        double xi[D], xbi[D], ni[D], nhat[D], tau_hat[D], pT[D], vi[D], gri[D], nrm[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = fluid ? s.x[i * D + a] : 0.0;
            xbi[a] = (FR && fluid) ? ax.xb[i * D + a] : 0.0;
            ni[a] = fluid ? s.norm[i * D + a] : 0.0;
            vi[a] = fluid ? s.v[i * D + a] : 0.0;
            gri[a] = fluid ? s.gradRho[i * D + a] : 0.0;
            nrm[a] = 0.0;
            tau_hat[a] = 0.0;
        }
        normalize<D>(ni, nhat);
        if (D == 2) {
            double tau[D];
            tau[0] = ni[1 % D];
            tau[1 % D] = -ni[0];
            normalize<D>(tau, tau_hat);
        }
#pragma unroll
        for (int a = 0; a < D; ++a) pT[a] = xi[a] + h133 * nhat[a];
        const int32_t bi = fluid ? s.b[i] : 0;
        const double lam_nb = fluid ? s.lam_nb[i] : 1.0;
        const double lami = fluid ? s.lam[i] : 1.0;
        const double rhoi = fluid ? s.rho[i] : 1.0;
        const bool testing = with_surf && fluid && bi >= FJSPH_PIPE && lam_nb >= 0.2 && lam_nb < 0.75;
        bool found = false;
        const bool high = lami > 0.7;
        double av[D], rd = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) av[a] = 0.0;
        const T Bg = T(c.B * c.gam), cfac = T(sqrt(c.B * c.gam / rhoi)), Ht = T(c.H);
        const T eps2 = T(0.0001 * c.H_sq), alpha = T(c.visc_alpha);
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
#pragma unroll
                for (int a = 0; a < D; ++a) tp[a][t] = s.x[(int64_t)sl * D + a];
                tvol[t] = ax.vol[sl];
                if constexpr (FR) tsl[t] = sl;
                if constexpr (with_surf) tlam[t] = s.lam[sl];
                if constexpr (with_diss) {
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        tv[a][t] = s.v[(int64_t)sl * D + a];
                        tg_[a][t] = s.gradRho[(int64_t)sl * D + a];
                    }
                    trho[t] = s.rho[sl];
                    tir[t] = ax.invrho[sl];
                    tm[t] = s.m[sl];
                    tnb[t] = s.b[sl] > FJSPH_PISTON;
                }
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                double xj[D], d2, r;
#pragma unroll
                for (int a = 0; a < D; ++a) xj[a] = tp[a][t];
                frozen_r<D, FR>(xi, xj, xbi, ax.xb, FR ? tsl[FR ? t : 0] : 0, d2, r);
                /* the outcome of the point-T test does not depend on the
                 * order of the pairs */
                if (testing && !found && marrone_pair<D>(xi, xj, r, nhat, tau_hat, pT, h133, r_lim)) found = true;
                double W, fg;
                wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
                if constexpr (with_surf) {
                    const double f = tvol[t] * (high ? (tlam[t] - lami) : tlam[t]) * fg;
#pragma unroll
                    for (int a = 0; a < D; ++a) nrm[a] += f * (xi[a] - xj[a]);
                }
                if constexpr (with_diss) {
                    /* dissipation_terms pair term (src/Shifting.cpp:141-173) */
                    T Rji[D], Vji[D], g[D];
                    T fgt;
                    if (sizeof(T) == 8) fgt = T(fg);
                    else { T Wt; wendland<T>(T(r), Ht, T(inv_H), T(c.Wc), Wt, fgt); }
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        Rji[a] = T(xj[a] - xi[a]);
                        Vji[a] = T(tv[a][t] - vi[a]);
                        g[a] = fgt * Rji[a];
                    }
                    const T idist2 = fast_rcp<T>(T(d2) + eps2);
                    const double rhoj = trho[t];
                    const T volj = T(tvol[t]);
                    T rdotg = Rji[0] * g[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rdotg += Rji[a] * g[a];
                    const T drho = T(rhoj - rhoi);
                    if (tnb[t]) {
                        T vdotr = Vji[0] * Rji[0];
#pragma unroll
                        for (int a = 1; a < D; ++a) vdotr += Vji[a] * Rji[a];
                        if (!(vdotr > T(0))) {
                            const T muij = Ht * vdotr * idist2;
                            const T cbar = T(0.5) * (cfac + fast_sqrt<T>(Bg * T(tir[t])));
                            const T ff = T(tm[t]) * alpha * cbar * muij * fast_rcp<T>(T(0.5 * (rhoi + rhoj)));
#pragma unroll
                            for (int a = 0; a < D; ++a) av[a] += double(g[a] * ff);
                        }
                        T gsum = T(0);
#pragma unroll
                        for (int a = 0; a < D; ++a) gsum += T(gri[a] + tg_[a][t]) * Rji[a];
                        rd += double(volj * (drho + T(0.5) * gsum) * rdotg * idist2);
                    } else {
                        rd += double(volj * drho * rdotg * idist2);
                    }
                }
            FJ_TILE_ENTRIES_END(cu, nl)
        }
        if (active) surface1_out<D, MODE>(c, s, i, fluid, bi, lam_nb, lami, found, nrm, av, rd, norms_tmp, ax.nhat);
    }
}

/* Mixed-precision form of k_surface1_t: the pair terms in FP32 on
 * coordinates relative to the block's first target (positions, velocities,
 * build positions) and densities relative to its density, summed in FP32
 * over each LDS tile and the tile sums in FP64; 64 B of LDS per candidate
 * (80 B with build positions).
 *
 * The point-T test is decided in FP32 unless one of its comparisons lies
 * within a margin of its threshold (1e-3 h133 for lengths, 1e-3 for the
 * cosine; the FP32 coordinates are off by at most |x - xref| 2^-24, well
 * under 1e-5 h133 for any block); such pairs repeat the test in FP64 from
 * the global positions, so the surface flags are those of the FP64 test.
 * acos(nhat . m) < pi/4 is evaluated as cos(pi/4) < nhat . m <= 1. */
#ifndef FJ_SURF1_OCC
#define FJ_SURF1_OCC __attribute__((amdgpu_waves_per_eu(5, 8)))
#endif
template <int D, int MODE, bool FR>
__global__ __launch_bounds__(TPB) FJ_SURF1_OCC void k_surface1_m(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                    TileGrid tg, NbrList nl, double* norms_tmp)
{
    constexpr bool with_surf = MODE != 2;
    constexpr bool with_diss = MODE != 0;
    constexpr int TD = with_diss ? TILE : 1;
    __shared__ float4 tA[TILE];            /* x_j - xref, V_j            */
    __shared__ float4 tB[TILE];            /* v_j - vref, lambda_j       */
    __shared__ float4 tG[TD];              /* gradRho_j, rho_j - rref    */
    __shared__ float2 tE[TD];              /* sqrt(B gam / rho_j), m_j (negative: j is PISTON or BOUND) */
    __shared__ float4 tX[FR ? TILE : 1];   /* xb_j - xbref               */
    __shared__ int32_t tsl[with_surf ? TILE : 1];
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double xref[D], vref[D], xbref[D], rref;
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            vref[a] = s.v[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
        rref = s.rho[i0];
    }
    const double h133 = 1.33 * c.particle_step;
    const double r_lim = sqrt(2.0) * h133;
    const float h133f = float(h133), r_limf = float(r_lim);
    const float tol = float(1e-3 * h133), tolc = 1e-3f;
    const float cos45 = float(cos(kPi / 4.0));
    const float Ht = float(c.H), inv_H = float(1.0 / c.H), Wc = float(c.Wc);
    const float eps2 = float(0.0001 * c.H_sq), alpha = float(c.visc_alpha);
    const float rreff = float(rref);
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        double nrm[D], av[D], rd = 0.0;
        /* per-tile FP32 partial sums, added into the FP64 totals after
         * each tile (about 40 pairs per tile and target) */
        float fnrm[D], fav[D], frd = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) { fnrm[a] = 0.f; fav[a] = 0.f; }
        float xit[D], xlo[D], xbit[D], xblo[D], vit[D], vlo[D], gri[D], nhf[D], tauf[D], pTt[D];
        {
            double xi[D], ni[D], nhat[D], tau_hat[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                xi[a] = fluid ? s.x[i * D + a] : xref[a];
                ni[a] = fluid ? s.norm[i * D + a] : 0.0;
                split_rel(xi[a], xref[a], xit[a], xlo[a]);
                split_rel((FR && fluid) ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
                split_rel(fluid ? s.v[i * D + a] : vref[a], vref[a], vit[a], vlo[a]);
                gri[a] = fluid ? float(s.gradRho[i * D + a]) : 0.f;
                nrm[a] = 0.0;
                av[a] = 0.0;
                tau_hat[a] = 0.0;
            }
            normalize<D>(ni, nhat);
            if (D == 2) {
                double tau[D];
                tau[0] = ni[1 % D];
                tau[1 % D] = -ni[0];
                normalize<D>(tau, tau_hat);
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                pTt[a] = float(xi[a] + h133 * nhat[a] - xref[a]);
                nhf[a] = float(nhat[a]);
                tauf[a] = float(tau_hat[a]);
            }
        }
        const int32_t bi = fluid ? s.b[i] : 0;
        const double lam_nb = fluid ? s.lam_nb[i] : 1.0;
        const double lami = fluid ? s.lam[i] : 1.0;
        const double rhoi = fluid ? s.rho[i] : rref;
        const bool testing = with_surf && fluid && bi >= FJSPH_PIPE && lam_nb >= 0.2 && lam_nb < 0.75;
        bool found = false;
        /* pairs whose FP32 point-T test is undecided (decided in FP64 by
         * k_surface1_resolve, so no FP64 call sits in the pair loop) */
        int npend = 0;
        const bool high = lami > 0.7;
        const float lamf = float(lami);
        float drhoi, rlo;
        split_rel(rhoi, rref, drhoi, rlo);
        const float rhoif = float(rhoi);
        const float cfac = float(sqrt(c.B * c.gam / rhoi));
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, va[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    if constexpr (with_diss) va[a] = float(s.v[(int64_t)sl * D + a] - vref[a]);
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                /* the slot's FP32 records: {gradRho, lambda} from dSPH_PreStep,
                 * {V, sqrt(B gamma / rho), +-m} from the aux refresh */
                const float4 r1 = ax.fs1[sl], r2 = ax.fs2[sl];
                tA[t] = make_float4(pa[0], pa[1], pa[2], r2.x);
                tB[t] = make_float4(va[0], va[1], va[2], with_surf ? r1.w : 0.f);
                if constexpr (FR) tX[t] = make_float4(xa[0], xa[1], xa[2], 0.f);
                if constexpr (with_surf) tsl[t] = sl;
                if constexpr (with_diss) {
                    tG[t] = make_float4(r1.x, r1.y, r1.z, float(s.rho[sl] - rref));
                    tE[t] = make_float2(r2.y, r2.z);
                }
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = tA[t];
                const float pj[3] = {A.x, A.y, A.z};
                float Rji[D], rr;
#pragma unroll
                for (int a = 0; a < D; ++a) Rji[a] = (pj[a] - xit[a]) - xlo[a];
                if constexpr (FR) {
                    const float4 X = tX[FR ? t : 0];
                    const float xbj[3] = {X.x, X.y, X.z};
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rji[0] * Rji[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rji[a] * Rji[a];
                }
                const float r = fast_sqrt<float>(rr);
                float W, fg;
                wendland<float>(r, Ht, inv_H, Wc, W, fg);
                const float4 B = tB[t];
                if constexpr (with_surf) {
                    if (testing && !found) {
                        /* point-T test (src/Geometry.cpp:43-98) in FP32 unless
                         * within the margin of a threshold */
                        float xjT[D];
#pragma unroll
                        for (int a = 0; a < D; ++a) xjT[a] = pj[a] - pTt[a];
                        const float dr = r - r_limf;
                        float q;
                        bool sure = fabsf(dr) >= tol;
                        if (dr >= 0.f) {
                            float z = xjT[0] * xjT[0];
#pragma unroll
                            for (int a = 1; a < D; ++a) z += xjT[a] * xjT[a];
                            q = h133f - fast_sqrt<float>(z);
                            sure = sure && fabsf(q) >= tol;
                        } else if (D == 2) {
                            float z1 = nhf[0] * xjT[0], z2 = tauf[0] * xjT[0];
#pragma unroll
                            for (int a = 1; a < D; ++a) { z1 += nhf[a] * xjT[a]; z2 += tauf[a] * xjT[a]; }
                            q = h133f - (fabsf(z1) + fabsf(z2));
                            sure = sure && fabsf(q) >= tol;
                        } else {
                            float z = nhf[0] * Rji[0];
#pragma unroll
                            for (int a = 1; a < D; ++a) z += nhf[a] * Rji[a];
                            /* acos of a cosine past 1 (moved positions,
                             * frozen r) is NaN: no hit.  The hardware
                             * reciprocal (1 ulp) sits far inside tolc; r = 0
                             * gives NaN either way, which is undecided */
                            const float cq = z * fast_rcp<float>(r);
                            q = (cq - cos45) * (1.f - cq);
                            sure = sure && fabsf(cq - cos45) >= tolc && fabsf(1.f - cq) >= tolc;
                        }
                        if (sure) found = q > 0.f;
                        else {
                            /* rare: the slot goes straight to memory */
                            if (npend < kSurfPendMax)
                                reinterpret_cast<int32_t*>(ax.spend)[i * kSurfPendMax + npend] = tsl[with_surf ? t : 0];
                            ++npend;
                        }
                    }
                    const float f = A.w * (high ? (B.w - lamf) : B.w) * fg;
#pragma unroll
                    for (int a = 0; a < D; ++a) fnrm[a] -= f * Rji[a];
                }
                if constexpr (with_diss) {
                    /* dissipation_terms pair term (src/Shifting.cpp:141-173) */
                    const float4 G = tG[t];
                    const float2 E = tE[t];
                    const float vj[3] = {B.x, B.y, B.z}, grj[3] = {G.x, G.y, G.z};
                    float g[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) g[a] = fg * Rji[a];
                    const float idist2 = fast_rcp<float>(rr + eps2);
                    float rdotg = Rji[0] * g[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rdotg += Rji[a] * g[a];
                    const float drho = (G.w - drhoi) - rlo;
                    if (E.y > 0.f) {
                        float vdotr = ((vj[0] - vit[0]) - vlo[0]) * Rji[0];
#pragma unroll
                        for (int a = 1; a < D; ++a) vdotr += ((vj[a] - vit[a]) - vlo[a]) * Rji[a];
                        if (!(vdotr > 0.f)) {
                            const float muij = Ht * vdotr * idist2;
                            const float cbar = 0.5f * (cfac + E.x);
                            const float ff =
                                E.y * alpha * cbar * muij * fast_rcp<float>(0.5f * (rhoif + (rreff + G.w)));
#pragma unroll
                            for (int a = 0; a < D; ++a) fav[a] += g[a] * ff;
                        }
                        float gsum = (gri[0] + grj[0]) * Rji[0];
#pragma unroll
                        for (int a = 1; a < D; ++a) gsum += (gri[a] + grj[a]) * Rji[a];
                        frd += A.w * (drho + 0.5f * gsum) * rdotg * idist2;
                    } else {
                        frd += A.w * drho * rdotg * idist2;
                    }
                }
            FJ_TILE_ENTRIES_END(cu, nl)
#pragma unroll
            for (int a = 0; a < D; ++a) {
                nrm[a] += double(fnrm[a]);
                av[a] += double(fav[a]);
                fnrm[a] = 0.f;
                fav[a] = 0.f;
            }
            rd += double(frd);
            frd = 0.f;
        }
        if (active) surface1_out<D, MODE>(c, s, i, fluid, bi, lam_nb, lami, found, nrm, av, rd, norms_tmp, ax.nhat);
        if constexpr (with_surf) {
            /* undecided: k_surface1_resolve sets the flag from the FP64 test */
            if (testing && !found && npend > 0) {
                s.surf[i] = kSurfUndecided + min(npend, kSurfPendMax + 1);
                ax.sund[1 + atomicAdd(ax.sund, 1)] = int32_t(i);
            }
        }
    }
}

__global__ void k_clear_count(int32_t* count) { *count = 0; }

/* Surface flag of a target that k_surface1_m left undecided (an FP32
 * point-T comparison within the margin of its threshold, and no sure hit):
 * the FP64 test of the undecided pairs (the others are surely out), or of
 * every pair when there were more than four (a sure FP32 decision equals the
 * FP64 one, so the OR is that of the FP64 test); src/Geometry.cpp:100-143
 * as in surface1_out.  One wave per undecided target: its lanes take the
 * pending slots, or split the list rows, and the flag is the OR over the
 * wave, so the dependent loads of a list walk run 64 at a time. */
template <int D, bool FR>
__global__ __launch_bounds__(TPB) void k_surface1_resolve(Consts c, DevState s, Aux ax, TileGrid tg, NbrList nl)
{
    const int32_t cnt = ax.sund[0];
    const int lane = threadIdx.x & 63;
    const int64_t nwave = int64_t(gridDim.x) * (TPB / 64);
    for (int64_t w = (blockIdx.x * (int64_t)TPB + threadIdx.x) >> 6; w < cnt; w += nwave) {
        const int64_t i = ax.sund[1 + w];
        const int np = s.surf[i] - kSurfUndecided;
        bool found = false;
        if (np <= kSurfPendMax) {
            if (lane < np) {
                const int4 pend = ax.spend[i];
                const int32_t sl = lane == 0 ? pend.x : lane == 1 ? pend.y : lane == 2 ? pend.z : pend.w;
                found = marrone_exact<D, FR>(c, s, ax, i, sl);
            }
        } else {
            int32_t seg_a[MAXSEG], pre[MAXSEG + 1];
            int nseg, p0, p1;
            union_of_block<D>(tg, block_of_slot(tg, i), seg_a, pre, nseg, p0, p1);
            const int NT = n_tiles(pre[nseg]);
            const int64_t p = tg.pos[i];
            const int nrows = min(nl.count[i], nl.cap >> 2);
            for (int r = lane; r < nrows && !found; r += 64) {
                const uint64_t u = nl.lu4[(int64_t)r * nl.npad + p];
                for (int q = 0; q < 4 && !found; ++q) {
                    const uint32_t uq = uint32_t(u >> (16 * q)) & 0xFFFFu;
                    if (uq != NO_ENTRY)
                        found = marrone_exact<D, FR>(c, s, ax, i, union_slot(tg, seg_a, pre, nseg, entry_union_index(uq, NT)));
                }
            }
        }
        found = __any(found);
        if (lane == 0) s.surf[i] = found ? 0 : 1;
    }
}

template <int D, typename T, int MODE>
static void launch_surface1_m(const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                              int64_t n, int r_from_x, double* norms_tmp, hipStream_t st)
{
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if constexpr (sizeof(T) == 4) {
        if constexpr (MODE != 2) hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(1), 0, st, a.sund);
        if (r_from_x)
            hipLaunchKernelGGL((k_surface1_m<D, MODE, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else
            hipLaunchKernelGGL((k_surface1_m<D, MODE, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        if constexpr (MODE != 2) {
            const dim3 gs((unsigned)std::min<int64_t>(grid_of(n), 1024)), bs(TPB);
            if (r_from_x) hipLaunchKernelGGL((k_surface1_resolve<D, false>), gs, bs, 0, st, c, s, a, tg, nl);
            else hipLaunchKernelGGL((k_surface1_resolve<D, true>), gs, bs, 0, st, c, s, a, tg, nl);
        }
    } else {
        if (r_from_x)
            hipLaunchKernelGGL((k_surface1_t<D, T, MODE, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else
            hipLaunchKernelGGL((k_surface1_t<D, T, MODE, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
    }
}

template <int D, typename T>
static void launch_surface1_d(const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                              int64_t n, int r_from_x, double* norms_tmp, int mode, hipStream_t st)
{
    if (mode == 0) launch_surface1_m<D, T, 0>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, st);
    else if (mode == 1) launch_surface1_m<D, T, 1>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, st);
    else launch_surface1_m<D, T, 2>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, st);
}

void launch_surface1(int dim, int prec, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg,
                     NbrList nl, int64_t n, int r_from_x, double* norms_tmp, int mode, hipStream_t st)
{
    const bool f32 = prec == FJSPH_PREC_MIXED;
    if (dim == 2) {
        if (f32) launch_surface1_d<2, float>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, mode, st);
        else launch_surface1_d<2, double>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, mode, st);
    } else {
        if (f32) launch_surface1_d<3, float>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, mode, st);
        else launch_surface1_d<3, double>(c, s, a, part_id, tg, nl, n, r_from_x, norms_tmp, mode, st);
    }
}

/* Curvature, Gissler occlusion and surfzone (src/Geometry.cpp:145-277). */
template <int D, bool FR>
__global__ __launch_bounds__(TPB) void k_surface2_t(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                    TileGrid tg, NbrList nl, const double* __restrict__ norms_tmp)
{
    __shared__ double tp[D][TILE], tn[D][TILE];
    __shared__ double tvol[TILE];
    __shared__ int32_t tsurf[TILE], tsl[FR ? TILE : 1];
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    const double inv_H = 1.0 / c.H;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        double xi[D], xbi[D], ni[D], Vd[D], Li[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = fluid ? s.x[i * D + a] : 0.0;
            xbi[a] = (FR && fluid) ? ax.xb[i * D + a] : 0.0;
            ni[a] = fluid ? norms_tmp[i * D + a] : 0.0;
            Vd[a] = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) Li[a][b2] = fluid ? s.L[i * D * D + a * D + b2] : 0.0;
        }
        const int32_t bi = fluid ? s.b[i] : 0;
        const bool occl = fluid && (c.aero_case == FJSPH_GISSLER) && bi == FJSPH_FREE;
        if (occl && s.cellID[i] != -1) {
#pragma unroll
            for (int a = 0; a < D; ++a)
                Vd[a] = (c.aero_source == FJSPH_MESH_INFL ? s.cellV[i * D + a] : c.v_inf[a]) - s.v[i * D + a];
        }
        const double Vdn = sqrt(dotd<D>(Vd, Vd));
        const bool ni_nz = dotd<D>(ni, ni) > 0.0;
        double curve = 0.0, qmax = 0.0;
        int zone = (fluid && s.surf[i] == 1) ? 1 : 0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    tp[a][t] = s.x[(int64_t)sl * D + a];
                    tn[a][t] = norms_tmp[(int64_t)sl * D + a];
                }
                tvol[t] = ax.vol[sl];
                tsurf[t] = s.surf[sl];
                if constexpr (FR) tsl[t] = sl;
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                double xj[D], d2, r, Rij[D], nj[D];
#pragma unroll
                for (int a = 0; a < D; ++a) { xj[a] = tp[a][t]; Rij[a] = xj[a] - xi[a]; nj[a] = tn[a][t]; }
                frozen_r<D, FR>(xi, xj, xbi, ax.xb, FR ? tsl[FR ? t : 0] : 0, d2, r);
                if (ni_nz && dotd<D>(nj, nj) > 0.0) {
                    double W, fg;
                    wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
                    double tt = 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        double la = 0.0;
#pragma unroll
                        for (int b2 = 0; b2 < D; ++b2) la += Li[a][b2] * (nj[b2] - ni[b2]);
                        tt += la * (fg * Rij[a]);
                    }
                    curve += tvol[t] * tt;
                }
                if (occl) {
                    /* frac = -Rij.Vd / (|Vd| r): the max is taken over
                     * -Rij.Vd / r and divided by |Vd| once */
                    const double q = -dotd<D>(Rij, Vd) / r;
                    if (q > qmax) qmax = q;
                }
                if (tsurf[t] == 1) zone = 1;
            FJ_TILE_ENTRIES_END(cu, nl)
        }
        if (!fluid) continue;
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
}

/* Mixed-precision form of k_surface2_t: FP32 pair terms on coordinates
 * relative to the block's first target; the curvature summed in FP32 per
 * LDS tile and in FP64 over the tiles; 32 B of LDS per candidate (48 B
 * with build positions).  The surface-zone flag is exact; a zero normal
 * stays exactly zero in FP32. */
#ifndef FJ_SURF2_OCC
#define FJ_SURF2_OCC __attribute__((amdgpu_waves_per_eu(5, 8)))
#endif
template <int D, bool FR>
__global__ __launch_bounds__(TPB) FJ_SURF2_OCC void k_surface2_m(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                    TileGrid tg, NbrList nl, const double* __restrict__ norms_tmp)
{
    __shared__ float4 tA[TILE];            /* x_j - xref, V_j */
    __shared__ float4 tN[TILE];            /* surface normal of j (Detect_Surface loop 1), surf_j */
    __shared__ float4 tX[FR ? TILE : 1];   /* xb_j - xbref */
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double xref[D], xbref[D];
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
    }
    const float Ht = float(c.H), inv_Hf = float(1.0 / c.H), Wcf = float(c.Wc);
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        float xit[D], xlo[D], xbit[D], xblo[D], ni[D], Vd[D], Li[D][D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            split_rel(fluid ? s.x[i * D + a] : xref[a], xref[a], xit[a], xlo[a]);
            split_rel((FR && fluid) ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
            ni[a] = fluid ? float(norms_tmp[i * D + a]) : 0.f;
            Vd[a] = 0.f;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) Li[a][b2] = fluid ? float(s.L[i * D * D + a * D + b2]) : 0.f;
        }
        const int32_t bi = fluid ? s.b[i] : 0;
        const bool occl = fluid && (c.aero_case == FJSPH_GISSLER) && bi == FJSPH_FREE;
        double Vd64[D];
#pragma unroll
        for (int a = 0; a < D; ++a) Vd64[a] = 0.0;
        if (occl && s.cellID[i] != -1) {
#pragma unroll
            for (int a = 0; a < D; ++a) {
                Vd64[a] = (c.aero_source == FJSPH_MESH_INFL ? s.cellV[i * D + a] : c.v_inf[a]) - s.v[i * D + a];
                Vd[a] = float(Vd64[a]);
            }
        }
        bool ni_nz = false;
#pragma unroll
        for (int a = 0; a < D; ++a) ni_nz = ni_nz || ni[a] != 0.f;
        double curve = 0.0;
        float fcurve = 0.f, qmax = 0.f;
        bool zone = fluid && s.surf[i] == 1;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, na[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    na[a] = float(norms_tmp[(int64_t)sl * D + a]);
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                tA[t] = make_float4(pa[0], pa[1], pa[2], float(ax.vol[sl]));
                tN[t] = make_float4(na[0], na[1], na[2], s.surf[sl] == 1 ? 1.f : 0.f);
                if constexpr (FR) tX[t] = make_float4(xa[0], xa[1], xa[2], 0.f);
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = tA[t], N = tN[t];
                const float pj[3] = {A.x, A.y, A.z}, nj[3] = {N.x, N.y, N.z};
                float Rij[D], rr;
#pragma unroll
                for (int a = 0; a < D; ++a) Rij[a] = (pj[a] - xit[a]) - xlo[a];
                if constexpr (FR) {
                    const float4 X = tX[FR ? t : 0];
                    const float xbj[3] = {X.x, X.y, X.z};
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rij[0] * Rij[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rij[a] * Rij[a];
                }
                const float r = fast_sqrt<float>(rr);
                bool nj_nz = false;
#pragma unroll
                for (int a = 0; a < D; ++a) nj_nz = nj_nz || nj[a] != 0.f;
                if (ni_nz && nj_nz) {
                    float W, fg;
                    wendland<float>(r, Ht, inv_Hf, Wcf, W, fg);
                    float tt = 0.f;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        float la = 0.f;
#pragma unroll
                        for (int b2 = 0; b2 < D; ++b2) la += Li[a][b2] * (nj[b2] - ni[b2]);
                        tt += la * (fg * Rij[a]);
                    }
                    fcurve += A.w * tt;
                }
                if (occl) {
                    float q = Rij[0] * Vd[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) q += Rij[a] * Vd[a];
                    q = -q / r;
                    if (q > qmax) qmax = q;
                }
                zone = zone || N.w != 0.f;
            FJ_TILE_ENTRIES_END(cu, nl)
            curve += double(fcurve);
            fcurve = 0.f;
        }
        if (!fluid) continue;
        const double Vdn = sqrt(dotd<D>(Vd64, Vd64));
        const double wo = (occl && Vdn > 0.0) ? double(qmax) / Vdn : 0.0;
        s.woccl[i] = s.lam_nb[i] < c.lam_cutoff ? fmax(0.0, fmin(wo, 1.0)) : 1.0;
        double n64[D], nh[D];
        load_vec<D>(norms_tmp, i, n64);
        normalize<D>(n64, nh);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            s.norm[i * D + a] = n64[a];
            ax.nhat[i * D + a] = nh[a];
        }
        s.curve[i] = curve;
        s.norm_curve[i] = c.dx * curve;
        s.pDist[i] = s.lam[i];
        if (c.ale) s.surfzone[i] = zone ? 1 : 0;
    }
}

void launch_surface2(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                     int64_t n, int r_from_x, const double* norms_tmp, hipStream_t st)
{
    (void)n;
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if (nl.f32) {
        if (dim == 2) {
            if (r_from_x) hipLaunchKernelGGL((k_surface2_m<2, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
            else hipLaunchKernelGGL((k_surface2_m<2, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        } else {
            if (r_from_x) hipLaunchKernelGGL((k_surface2_m<3, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
            else hipLaunchKernelGGL((k_surface2_m<3, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        }
        return;
    }
    if (dim == 2) {
        if (r_from_x) hipLaunchKernelGGL((k_surface2_t<2, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else hipLaunchKernelGGL((k_surface2_t<2, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
    } else {
        if (r_from_x) hipLaunchKernelGGL((k_surface2_t<3, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else hipLaunchKernelGGL((k_surface2_t<3, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
    }
}

/* ------------------------------------------------------------ particle_shift */

/* Per-slot result of particle_shift (src/Shifting.cpp:254-289) from the
 * pair sums: du = sum_j (1 + 0.2 (W/W_dx)^4) V_j W'(r)/r R_ji, maxU2 =
 * max_j |v_j - v_i|^2, dmin = min_j n_i . n_j over the non-boundary j. */
template <int D>
__device__ __forceinline__ void shift_out(const Consts& c, DevState& s, int64_t i, bool work, double lam_nb,
                                          const double* vi, const double* nhat, double* du, double maxU2, double dmin,
                                          int32_t zone)
{
    double out[D];
    if (!work) {
#pragma unroll
        for (int a = 0; a < D; ++a) out[a] = 0.0;
    } else {
        const double wo = dmin <= 1.0 ? acos(dmin) : 0.0;
        const double fac = -2.0 * c.H * sqrt(dotd<D>(vi, vi));
#pragma unroll
        for (int a = 0; a < D; ++a) du[a] *= fac;
        double dun[D];
        normalize<D>(du, dun);
        const double mag = fmin(sqrt(dotd<D>(du, du)), fmin(sqrt(maxU2) / 2.0, c.max_shift_vel));
#pragma unroll
        for (int a = 0; a < D; ++a) du[a] = mag * dun[a];
        if (zone == 0 && lam_nb > 0.55) {
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
    }
#pragma unroll
    for (int a = 0; a < D; ++a) s.vPert[i * D + a] = out[a];
}

/* src/Shifting.cpp:189-290.  max_j acos(n_i.n_j) is taken as acos(min_j
 * n_i.n_j) over the dot products acos accepts ([-1, 1]): NaN angles never
 * win the reference's max and acos is decreasing, so the result is the same.
 * max_j |v_j - v_i| is taken on squared norms. */
template <int D, bool FR>
__global__ __launch_bounds__(TPB) void k_shift_t(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                 TileGrid tg, NbrList nl)
{
    __shared__ double tp[D][TILE], tv[D][TILE], tn[D][TILE];
    __shared__ double tvol[TILE];
    __shared__ int32_t tnb[TILE], tsl[FR ? TILE : 1];
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    const double inv_H = 1.0 / c.H;
    const double inv_Wdx = 1.0 / c.W_dx;
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        const double lam_nb = fluid ? s.lam_nb[i] : 0.0;
        const bool work = fluid && !(lam_nb < 0.55 || s.b[i] == FJSPH_BUFFER);
        double xi[D], xbi[D], vi[D], nhat[D], du[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = work ? s.x[i * D + a] : 0.0;
            xbi[a] = (FR && work) ? ax.xb[i * D + a] : 0.0;
            vi[a] = work ? s.v[i * D + a] : 0.0;
            nhat[a] = work ? ax.nhat[i * D + a] : 0.0;
            du[a] = 0.0;
        }
        double maxU2 = 0.0, dmin = 2.0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, work);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    tp[a][t] = s.x[(int64_t)sl * D + a];
                    tv[a][t] = s.v[(int64_t)sl * D + a];
                    tn[a][t] = ax.nhat[(int64_t)sl * D + a];
                }
                tvol[t] = ax.vol[sl];
                tnb[t] = s.b[sl] > FJSPH_PISTON;
                if constexpr (FR) tsl[t] = sl;
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                double xj[D], d2, r;
#pragma unroll
                for (int a = 0; a < D; ++a) xj[a] = tp[a][t];
                frozen_r<D, FR>(xi, xj, xbi, ax.xb, FR ? tsl[FR ? t : 0] : 0, d2, r);
                double W, fg;
                wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
                const double w = W * inv_Wdx;
                const double w2 = w * w;
                const double f = (1.0 + 0.2 * (w2 * w2)) * fg * tvol[t];
                double dv2 = 0.0;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    du[a] += f * (xj[a] - xi[a]);
                    const double dv = tv[a][t] - vi[a];
                    dv2 += dv * dv;
                }
                if (tnb[t]) {
                    double d = 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) d += nhat[a] * tn[a][t];
                    if (d >= -1.0 && d <= 1.0 && d < dmin) dmin = d;
                }
                if (dv2 > maxU2) maxU2 = dv2;
            FJ_TILE_ENTRIES_END(cu, nl)
        }
        if (fluid) shift_out<D>(c, s, i, work, lam_nb, vi, nhat, du, maxU2, dmin, s.surfzone[i]);
    }
}

/* Mixed-precision form of k_shift_t: FP32 pair terms on positions and
 * velocities relative to the block's first target, the shift sum in FP32
 * per LDS tile and in FP64 over the tiles; the normal products n_i . n_j
 * stay in FP64 (they decide between the projected and the zero shift);
 * 56 B of LDS per candidate (72 B with build positions). */
#ifndef FJ_SHIFT_OCC
#define FJ_SHIFT_OCC __attribute__((amdgpu_waves_per_eu(5, 8)))
#endif
template <int D, bool FR>
__global__ __launch_bounds__(TPB) FJ_SHIFT_OCC void k_shift_m(Consts c, DevState s, Aux ax, const int32_t* __restrict__ part_id,
                                                 TileGrid tg, NbrList nl)
{
    __shared__ float4 tA[TILE];            /* x_j - xref, +V_j (-V_j: j is PISTON or BOUND) */
    __shared__ float4 tV[TILE];            /* v_j - vref */
    __shared__ double tn[D][TILE];         /* normalised normal of j */
    __shared__ float4 tX[FR ? TILE : 1];   /* xb_j - xbref */
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double xref[D], vref[D], xbref[D];
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            vref[a] = s.v[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
    }
    const float Ht = float(c.H), inv_Hf = float(1.0 / c.H), Wcf = float(c.Wc), inv_Wdx = float(1.0 / c.W_dx);
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        const double lam_nb = fluid ? s.lam_nb[i] : 0.0;
        const bool work = fluid && !(lam_nb < 0.55 || s.b[i] == FJSPH_BUFFER);
        double nhat[D], du[D];
        float xit[D], xlo[D], xbit[D], xblo[D], vit[D], vlo[D], fdu[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            split_rel(work ? s.x[i * D + a] : xref[a], xref[a], xit[a], xlo[a]);
            split_rel((FR && work) ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
            split_rel(work ? s.v[i * D + a] : vref[a], vref[a], vit[a], vlo[a]);
            nhat[a] = work ? ax.nhat[i * D + a] : 0.0;
            du[a] = 0.0;
            fdu[a] = 0.f;
        }
        float maxU2 = 0.f;
        double dmin = 2.0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, work);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, va[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    va[a] = float(s.v[(int64_t)sl * D + a] - vref[a]);
                    tn[a][t] = ax.nhat[(int64_t)sl * D + a];
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                const float v = float(ax.vol[sl]);
                tA[t] = make_float4(pa[0], pa[1], pa[2], s.b[sl] > FJSPH_PISTON ? v : -v);
                tV[t] = make_float4(va[0], va[1], va[2], 0.f);
                if constexpr (FR) tX[t] = make_float4(xa[0], xa[1], xa[2], 0.f);
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = tA[t], V = tV[t];
                const float pj[3] = {A.x, A.y, A.z}, vj[3] = {V.x, V.y, V.z};
                float Rji[D], rr;
#pragma unroll
                for (int a = 0; a < D; ++a) Rji[a] = (pj[a] - xit[a]) - xlo[a];
                if constexpr (FR) {
                    const float4 X = tX[FR ? t : 0];
                    const float xbj[3] = {X.x, X.y, X.z};
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rji[0] * Rji[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rji[a] * Rji[a];
                }
                float W, fg;
                wendland<float>(fast_sqrt<float>(rr), Ht, inv_Hf, Wcf, W, fg);
                const float w = W * inv_Wdx;
                const float w2 = w * w;
                const float f = (1.f + 0.2f * (w2 * w2)) * fg * fabsf(A.w);
                float dv2 = 0.f;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    fdu[a] += f * Rji[a];
                    const float dv = (vj[a] - vit[a]) - vlo[a];
                    dv2 += dv * dv;
                }
                if (A.w > 0.f) {
                    double d = 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) d += nhat[a] * tn[a][t];
                    if (d >= -1.0 && d <= 1.0 && d < dmin) dmin = d;
                }
                maxU2 = fmaxf(maxU2, dv2);
            FJ_TILE_ENTRIES_END(cu, nl)
#pragma unroll
            for (int a = 0; a < D; ++a) {
                du[a] += double(fdu[a]);
                fdu[a] = 0.f;
            }
        }
        if (!fluid) continue;
        double vi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) vi[a] = work ? s.v[i * D + a] : 0.0;
        shift_out<D>(c, s, i, work, lam_nb, vi, nhat, du, double(maxU2), dmin, s.surfzone[i]);
    }
}

void launch_shift(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                  int64_t n, int r_from_x, hipStream_t st)
{
    (void)n;
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if (nl.f32) {
        if (dim == 2) {
            if (r_from_x) hipLaunchKernelGGL((k_shift_m<2, false>), g, b, 0, st, c, s, a, part_id, tg, nl);
            else hipLaunchKernelGGL((k_shift_m<2, true>), g, b, 0, st, c, s, a, part_id, tg, nl);
        } else {
            if (r_from_x) hipLaunchKernelGGL((k_shift_m<3, false>), g, b, 0, st, c, s, a, part_id, tg, nl);
            else hipLaunchKernelGGL((k_shift_m<3, true>), g, b, 0, st, c, s, a, part_id, tg, nl);
        }
        return;
    }
    if (dim == 2) {
        if (r_from_x) hipLaunchKernelGGL((k_shift_t<2, false>), g, b, 0, st, c, s, a, part_id, tg, nl);
        else hipLaunchKernelGGL((k_shift_t<2, true>), g, b, 0, st, c, s, a, part_id, tg, nl);
    } else {
        if (r_from_x) hipLaunchKernelGGL((k_shift_t<3, false>), g, b, 0, st, c, s, a, part_id, tg, nl);
        else hipLaunchKernelGGL((k_shift_t<3, true>), g, b, 0, st, c, s, a, part_id, tg, nl);
    }
}

/* Detect_Surface loops 2-3 (k_surface2_m) and particle_shift (k_shift_m)
 * in one staged pass, for ALE builds in mixed precision: the shifting pass
 * reads of its neighbours only what Detect_Surface loop 1 and dSPH_PreStep
 * produced (positions, velocities, volumes and the unit normal, the
 * normalised loop-1 normal that loop 1 writes to Aux::nhat) and of its
 * target the surface-zone flag that loops 2-3 form in the same walk, so one
 * staging and one list walk serve both (src/Geometry.cpp:145-277,
 * src/Shifting.cpp:189-290).  The pair terms and every decision are those
 * of the two separate kernels, which the per-pass entry points keep; the
 * FP32 normals of the curvature sum are the unit normals (the loop-1
 * normals are unit vectors or zero, so they differ in the last bit at
 * most).  72 B of LDS per candidate (88 B with build positions). */
#ifndef FJ_SURF2SHIFT_OCC
#define FJ_SURF2SHIFT_OCC __attribute__((amdgpu_waves_per_eu(4, 8)))
#endif
template <int D, bool FR>
__global__ __launch_bounds__(TPB) FJ_SURF2SHIFT_OCC void k_surf2_shift_m(Consts c, DevState s, Aux ax,
                                                                   const int32_t* __restrict__ part_id, TileGrid tg,
                                                                   NbrList nl, const double* __restrict__ norms_tmp)
{
    __shared__ float4 tA[TILE];            /* x_j - xref, +V_j (-V_j: j is PISTON or BOUND) */
    __shared__ float4 tN[TILE];            /* surface normal of j (Detect_Surface loop 1), surf_j */
    __shared__ float4 tV[TILE];            /* v_j - vref */
    __shared__ int32_t tsl[TILE];          /* particle slot of j (FP64 normal of an undecided angle test) */
    __shared__ float4 tX[FR ? TILE : 1];   /* xb_j - xbref */
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double xref[D], vref[D], xbref[D];
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            vref[a] = s.v[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
    }
    const float Ht = float(c.H), inv_Hf = float(1.0 / c.H), Wcf = float(c.Wc), inv_Wdx = float(1.0 / c.W_dx);
    /* particle_shift's angle test (src/Shifting.cpp:241-243, 273),
     * acos(n_i . n_j) >= pi/12 for some fluid j whose dot lies in [-1, 1],
     * decided on the FP32 dot away from its two thresholds: the FP32 dot of
     * the rounded unit normals is within 4e-7 of the FP64 one, the margin
     * is 1e-5 */
    const double cos15 = cos(kPi / 12.0);
    const float ang_yes_lo = float(-1.0 + 1e-5), ang_yes_hi = float(cos15 - 1e-5), ang_no = float(cos15 + 1e-5);
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        const double lam_nb = fluid ? s.lam_nb[i] : 0.0;
        const bool work = fluid && !(lam_nb < 0.55 || s.b[i] == FJSPH_BUFFER);
        float xit[D], xlo[D], xbit[D], xblo[D], vit[D], vlo[D], ni[D], Vd[D], Li[D][D], fdu[D];
        double nhat[D], du[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            nhat[a] = fluid ? ax.nhat[i * D + a] : 0.0;
            ni[a] = float(nhat[a]);
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            split_rel(fluid ? s.x[i * D + a] : xref[a], xref[a], xit[a], xlo[a]);
            split_rel((FR && fluid) ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
            split_rel(work ? s.v[i * D + a] : vref[a], vref[a], vit[a], vlo[a]);
            Vd[a] = 0.f;
            du[a] = 0.0;
            fdu[a] = 0.f;
#pragma unroll
            for (int b2 = 0; b2 < D; ++b2) Li[a][b2] = fluid ? float(s.L[i * D * D + a * D + b2]) : 0.f;
        }
        const int32_t bi = fluid ? s.b[i] : 0;
        const bool occl = fluid && (c.aero_case == FJSPH_GISSLER) && bi == FJSPH_FREE;
        /* the gas velocity relative to the target, in FP32 for the pairs
         * (its FP64 norm is formed again after the walk) */
        const bool has_vd = occl && s.cellID[i] != -1;
        if (has_vd) {
#pragma unroll
            for (int a = 0; a < D; ++a)
                Vd[a] = float((c.aero_source == FJSPH_MESH_INFL ? s.cellV[i * D + a] : c.v_inf[a]) - s.v[i * D + a]);
        }
        bool ni_nz = false;
#pragma unroll
        for (int a = 0; a < D; ++a) ni_nz = ni_nz || ni[a] != 0.f;
        double curve = 0.0;
        float fcurve = 0.f, qmax = 0.f, maxU2 = 0.f;
        bool zone = fluid && s.surf[i] == 1, wide = false;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, va[3] = {0.f, 0.f, 0.f}, na[3] = {0.f, 0.f, 0.f},
                      xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    va[a] = float(s.v[(int64_t)sl * D + a] - vref[a]);
                    na[a] = float(ax.nhat[(int64_t)sl * D + a]);
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                tsl[t] = sl;
                const float4 r2 = ax.fs2[sl]; /* V, .., m negative for PISTON and BOUND */
                tA[t] = make_float4(pa[0], pa[1], pa[2], copysignf(r2.x, r2.z));
                tN[t] = make_float4(na[0], na[1], na[2], s.surf[sl] == 1 ? 1.f : 0.f);
                tV[t] = make_float4(va[0], va[1], va[2], 0.f);
                if constexpr (FR) tX[t] = make_float4(xa[0], xa[1], xa[2], 0.f);
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = tA[t], N = tN[t];
                const float pj[3] = {A.x, A.y, A.z}, nj[3] = {N.x, N.y, N.z};
                const float volj = fabsf(A.w);
                float Rij[D], rr;
#pragma unroll
                for (int a = 0; a < D; ++a) Rij[a] = (pj[a] - xit[a]) - xlo[a];
                if constexpr (FR) {
                    const float4 X = tX[FR ? t : 0];
                    const float xbj[3] = {X.x, X.y, X.z};
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rij[0] * Rij[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rij[a] * Rij[a];
                }
                const float r = fast_sqrt<float>(rr);
                float W, fg;
                wendland<float>(r, Ht, inv_Hf, Wcf, W, fg);
                /* Detect_Surface loops 2-3: curvature, occlusion, surface zone */
                bool nj_nz = false;
#pragma unroll
                for (int a = 0; a < D; ++a) nj_nz = nj_nz || nj[a] != 0.f;
                if (ni_nz && nj_nz) {
                    float tt = 0.f;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        float la = 0.f;
#pragma unroll
                        for (int b2 = 0; b2 < D; ++b2) la += Li[a][b2] * (nj[b2] - ni[b2]);
                        tt += la * (fg * Rij[a]);
                    }
                    fcurve += volj * tt;
                }
                if (occl) {
                    /* -R_ij . Vd / r with the hardware reciprocal (1 ulp) in
                     * place of the IEEE division */
                    float q = Rij[0] * Vd[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) q += Rij[a] * Vd[a];
                    q = -q * fast_rcp<float>(r);
                    if (q > qmax) qmax = q;
                }
                zone = zone || N.w != 0.f;
                /* particle_shift */
                if (work) {
                    const float4 V = tV[t];
                    const float vj[3] = {V.x, V.y, V.z};
                    const float w = W * inv_Wdx;
                    const float w2 = w * w;
                    const float f = (1.f + 0.2f * (w2 * w2)) * fg * volj;
                    float dv2 = 0.f;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        fdu[a] += f * Rij[a];
                        const float dv = (vj[a] - vit[a]) - vlo[a];
                        dv2 += dv * dv;
                    }
                    if (A.w > 0.f && !wide) {
                        float d = ni[0] * nj[0];
#pragma unroll
                        for (int a = 1; a < D; ++a) d += ni[a] * nj[a];
                        if (d >= ang_yes_lo && d <= ang_yes_hi) wide = true;
                        else if (!(d >= ang_no)) {
                            /* rare: within the margin of cos(pi/12) or -1 */
                            const int32_t sl = tsl[t];
                            double d64 = 0.0;
#pragma unroll
                            for (int a = 0; a < D; ++a) d64 += nhat[a] * ax.nhat[(int64_t)sl * D + a];
                            if (d64 >= -1.0 && d64 <= 1.0 && acos(d64) >= kPi / 12.0) wide = true;
                        }
                    }
                    maxU2 = fmaxf(maxU2, dv2);
                }
            FJ_TILE_ENTRIES_END(cu, nl)
            curve += double(fcurve);
            fcurve = 0.f;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                du[a] += double(fdu[a]);
                fdu[a] = 0.f;
            }
        }
        if (!fluid) continue;
        double Vd64[D];
#pragma unroll
        for (int a = 0; a < D; ++a)
            Vd64[a] = has_vd ? (c.aero_source == FJSPH_MESH_INFL ? s.cellV[i * D + a] : c.v_inf[a]) - s.v[i * D + a]
                             : 0.0;
        const double Vdn = sqrt(dotd<D>(Vd64, Vd64));
        const double wo = (occl && Vdn > 0.0) ? double(qmax) / Vdn : 0.0;
        s.woccl[i] = lam_nb < c.lam_cutoff ? fmax(0.0, fmin(wo, 1.0)) : 1.0;
#pragma unroll
        for (int a = 0; a < D; ++a) s.norm[i * D + a] = norms_tmp[i * D + a];
        s.curve[i] = curve;
        s.norm_curve[i] = c.dx * curve;
        s.pDist[i] = s.lam[i];
        s.surfzone[i] = zone ? 1 : 0;
        double vi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) vi[a] = work ? s.v[i * D + a] : 0.0;
        /* min dot 0 (acos = pi/2) when some angle reached pi/12, else 2
         * (no dot in range) */
        shift_out<D>(c, s, i, work, lam_nb, vi, nhat, du, double(maxU2), wide ? 0.0 : 2.0, zone ? 1 : 0);
    }
}

void launch_surface2_shift(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg,
                           NbrList nl, int64_t n, int r_from_x, const double* norms_tmp, hipStream_t st)
{
    (void)n;
    const dim3 g((unsigned)tg.nblocks), b(TPB);
    if (dim == 2) {
        if (r_from_x) hipLaunchKernelGGL((k_surf2_shift_m<2, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else hipLaunchKernelGGL((k_surf2_shift_m<2, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
    } else {
        if (r_from_x) hipLaunchKernelGGL((k_surf2_shift_m<3, false>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
        else hipLaunchKernelGGL((k_surf2_shift_m<3, true>), g, b, 0, st, c, s, a, part_id, tg, nl, norms_tmp);
    }
}

/* ------------------------------------------------------- boundary solvers */


/* The boundary passes run one wave per boundary slot: its 64 lanes split
 * the slot's list rows, so the dependent loads of a neighbour walk run 64
 * at a time instead of one after another in a single lane (a boundary
 * layer is a few thousand slots, each with ~200 neighbours), and the lane
 * sums meet in a fixed-order wave reduction. */
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* f(j, r) for every neighbour j of slot i, spread over the lanes of the
 * wave (all lanes call it with the same i); r is the distance of the last
 * build. */
template <int D, typename F>
__device__ __forceinline__ void walk_wave_r(const TileGrid& tg, const NbrList& nl, const double* __restrict__ xb,
                                            int64_t i, F&& f)
{
    const int lane = threadIdx.x & 63;
    int32_t seg_a[MAXSEG], pre[MAXSEG + 1];
    int nseg, p0, p1;
    union_of_block<D>(tg, block_of_slot(tg, i), seg_a, pre, nseg, p0, p1);
    const int NT = n_tiles(pre[nseg]);
    const int64_t p = tg.pos[i];
    const int nrows = min(nl.count[i], nl.cap >> 2);
    double xbi[D];
    load_vec<D>(xb, i, xbi);
    for (int r = lane; r < nrows; r += 64) {
        const uint64_t u = nl.lu4[(int64_t)r * nl.npad + p];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t uq = uint32_t(u >> (16 * q)) & 0xFFFFu;
            if (uq == NO_ENTRY) continue;
            const int32_t j = union_slot(tg, seg_a, pre, nseg, entry_union_index(uq, NT));
            double xbj[D];
            load_vec<D>(xb, j, xbj);
            f(j, sqrt(exact_dist2<D>(xbi, xbj)));
        }
    }
}

/* The slots of boundary block `block`, appended (count at out[0]). */
__global__ __launch_bounds__(TPB) void k_block_slots(const int32_t* __restrict__ blk, int64_t n, int32_t block,
                                                     int32_t* out)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i < n && blk[i] == block) out[1 + atomicAdd(out, 1)] = int32_t(i);
}

/* the boundary slot of this wave, or -1 past the list */
__device__ __forceinline__ int64_t wave_slot(const int32_t* __restrict__ slots)
{
    const int64_t w = (blockIdx.x * (int64_t)TPB + threadIdx.x) >> 6;
    return w < slots[0] ? int64_t(slots[1 + w]) : int64_t(-1);
}

/* Block velocity + no-slip (src/Newmark_Beta.cpp:69-102, src/Resid.cpp:162-186). */
template <int D>
__global__ __launch_bounds__(TPB) void k_bnd_vel(Consts c, DevState s, Aux ax, DevBlocks bl, int32_t block,
                                                 TileGrid tg, NbrList nl, const int32_t* __restrict__ slots)
{
    const int64_t i = wave_slot(slots);
    if (i < 0) return;
    double v[D];
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = bl.vel[block * 3 + a];
    if (bl.no_slip[block]) {
        const double inv_H = 1.0 / c.H;
        double vs[D];
#pragma unroll
        for (int a = 0; a < D; ++a) vs[a] = 0.0;
        double ks = 0.0;
        walk_wave_r<D>(tg, nl, ax.xb, i, [&](int32_t j, double r) {
            if (s.b[j] <= FJSPH_PISTON) return;
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            ks += W;
#pragma unroll
            for (int a = 0; a < D; ++a) vs[a] += s.v[(int64_t)j * D + a] * W;
        });
        ks = wave_sum(ks);
#pragma unroll
        for (int a = 0; a < D; ++a) vs[a] = wave_sum(vs[a]);
        if (ks > 0.0) {
#pragma unroll
            for (int a = 0; a < D; ++a) v[a] = 2.0 * v[a] - vs[a] / ks;
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < D; ++a) s.v[i * D + a] = v[a];
    }
}

/* Boundary solvers (src/Resid.cpp:21-186): DBC (acc = 0), Adami pressure
 * (Get_Boundary_Pressure), Ghost (boundary Rrho and the near-inlet flag). */
template <int D>
__global__ __launch_bounds__(TPB) void k_bnd_solve(Consts c, DevState s, Aux ax, DevBlocks bl, int32_t block,
                                                   TileGrid tg, NbrList nl, const int32_t* __restrict__ slots,
                                                   int32_t* near_inlet)
{
    const int64_t i = wave_slot(slots);
    if (i < 0) return;
    const bool lead = (threadIdx.x & 63) == 0;
    const int solver = bl.solver[block];
    const double inv_H = 1.0 / c.H;
    double xi[D];
    load_vec<D>(s.x, i, xi);
    if (solver == FJSPH_DBC) {
        if (lead) {
#pragma unroll
            for (int a = 0; a < D; ++a) s.acc[i * D + a] = 0.0;
        }
    } else if (solver == FJSPH_PRESSURE_G) {
        int near = 0;
        double ks = 0.0, pk = 0.0, ak[D];
#pragma unroll
        for (int a = 0; a < D; ++a) ak[a] = 0.0;
        walk_wave_r<D>(tg, nl, ax.xb, i, [&](int32_t j, double r) {
            if (s.b[j] <= FJSPH_PISTON) return;
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            const double kern = ax.vol[j] * W;
            ks += kern;
            pk += s.p[j] * kern;
            const double kr = kern * s.rho[j];
#pragma unroll
            for (int a = 0; a < D; ++a) ak[a] += kr * (xi[a] - s.x[(int64_t)j * D + a]);
            if (s.surfzone[j]) near = 1;
        });
        ks = wave_sum(ks);
        pk = wave_sum(pk);
#pragma unroll
        for (int a = 0; a < D; ++a) ak[a] = wave_sum(ak[a]);
        near = __any(near) ? 1 : 0;
        double pres = 0.0;
        if (ks > 0.0) {
            double ga = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) ga += (c.grav[a] - s.acc[i * D + a]) * ak[a];
            const double pp = (pk + ga) / ks;
            pres = near ? fmax(0.0, pp) : pp;
        }
        if (lead) {
            s.p[i] = pres;
            s.rho[i] = eos_density(c, pres);
        }
    } else if (solver == FJSPH_GHOST) {
        int ni = 1;
        double rr = 0.0, vi[D];
        load_vec<D>(s.v, i, vi);
        walk_wave_r<D>(tg, nl, ax.xb, i, [&](int32_t j, double r) {
            const int32_t bj = s.b[j];
            if (bj == FJSPH_PIPE || bj == FJSPH_FREE) ni = 0;
            double W, fg;
            wendland<double>(r, c.H, inv_H, c.Wc, W, fg);
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a)
                t += (s.v[(int64_t)j * D + a] - vi[a]) * (fg * (s.x[(int64_t)j * D + a] - xi[a]));
            rr -= ax.vol[j] * t;
        });
        rr = wave_sum(rr);
        ni = __any(ni == 0) ? 0 : 1;
        if (lead) {
            near_inlet[i] = ni;
            s.Rrho[i] = rr * s.rho[i];
        }
    }
}

void launch_boundary(int dim, const Consts& c, DevState s, Aux a, const int32_t* blk, DevBlocks bl, int32_t block,
                     TileGrid tg, NbrList nl, int64_t n, int64_t nb, int32_t* slots, int32_t* near_inlet,
                     hipStream_t st)
{
    /* the block's slots (at most nb, the boundary count), then one wave each */
    hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(1), 0, st, slots);
    hipLaunchKernelGGL(k_block_slots, dim3(grid_of(n)), dim3(TPB), 0, st, blk, n, block, slots);
    const dim3 g((unsigned)((nb + TPB / 64 - 1) / (TPB / 64))), b(TPB);
    if (dim == 2) {
        hipLaunchKernelGGL(k_bnd_vel<2>, g, b, 0, st, c, s, a, bl, block, tg, nl, slots);
        hipLaunchKernelGGL(k_bnd_solve<2>, g, b, 0, st, c, s, a, bl, block, tg, nl, slots, near_inlet);
    } else {
        hipLaunchKernelGGL(k_bnd_vel<3>, g, b, 0, st, c, s, a, bl, block, tg, nl, slots);
        hipLaunchKernelGGL(k_bnd_solve<3>, g, b, 0, st, c, s, a, bl, block, tg, nl, slots, near_inlet);
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


/* Eigen's normalized(): a zero vector is returned unchanged. */
template <int D>
__device__ void normalized_d(const double* a, double* out)
{
    const double z = dotd<D>(a, a);
    const double r = z > 0.0 ? sqrt(z) : 1.0;
#pragma unroll
    for (int k = 0; k < D; ++k) out[k] = z > 0.0 ? a[k] / r : a[k];
}

/* GetCd (src/Aero.h:10-18). */
__device__ inline double sphere_cd(double Re)
{
    return (1.0 + 0.197 * pow(Re, 0.63) + 2.6e-04 * pow(Re, 1.38)) * (24.0 / (Re + 0.00001));
}

/* induced_pressure (src/Aero.h:106-202) in FP64 for one particle: a
 * pressure coefficient blended from sphere, plate and bowl curves by the
 * angle between the surface normal and the gas velocity and by the
 * particle's scaled curvature, the induced pressure on the plate area along
 * the normal, a seventh-power-law skin friction on the tangential velocity,
 * blended with the droplet drag by the neighbourhood fraction. */
template <int D>
__device__ void induced_pressure(const Consts& c, const double* Vd, const double* nrm, double lam, double nneigh,
                                 double norm_curve, double cellP, double cellRho, double mass, double* out)
{
    double nn[D], vh[D];
    normalized_d<D>(nrm, nn);
    normalized_d<D>(Vd, vh);
    double cth = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) cth += -nn[k] * vh[k];
    const double theta = fabs(acos(cth));
    double Cp_s, Cp_p, Cp_b, Cp_tot;
    if (theta < 2.4455) {
        const double sn = sin(theta);
        Cp_s = 1.0 - (2.25) * (sn * sn);
    } else {
        Cp_s = 0.075;
    }
    if (theta < 1.570797) Cp_p = cos(theta);
    else if (theta < 1.9918) Cp_p = -pow(cos(6.0 * theta + 0.5 * kPi), 1.5);
    else if (theta < 2.0838) Cp_p = 5.5836 * theta - 11.5601;
    else Cp_p = 0.075;
    if (theta < 0.7854) Cp_b = 1.0;
    else if (theta < 1.570797) Cp_b = 0.5 * (cos(4.0 * theta - kPi) + 1.0);
    else Cp_b = 0.0;
    const double fac1 = 0.25, ifac1 = 1 / fac1;
    if (norm_curve < -fac1) {
        Cp_tot = Cp_b;
    } else if (norm_curve < 0.0) {
        const double frac = (norm_curve + fac1) * ifac1;
        Cp_tot = frac * Cp_b + (1.0 - frac) * Cp_p;
    } else if (norm_curve < fac1) {
        const double frac = norm_curve * ifac1;
        Cp_tot = frac * Cp_p + (1.0 - frac) * Cp_s;
    } else {
        Cp_tot = Cp_s;
    }
    const double sos2 = c.sos * c.sos;
    const double V2 = dotd<D>(Vd, Vd);
    const double Pi = 0.5 * V2 / sos2 * c.gamma_g * cellP * Cp_tot;
    const double Vn = sqrt(V2);
    const double Cdi = sphere_cd(cellRho * Vn * c.L / c.mu_g);
    double Vnorm = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) Vnorm += Vd[k] * nn[k];
    double Vpar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) Vpar[k] = Vd[k] - Vnorm * nn[k];
    const double Vpn = sqrt(dotd<D>(Vpar, Vpar));
    const double Cf = 0.027 / pow(cellRho * Vpn * c.L / c.mu_g + 1e-6, 1.0 / 7.0);
    const double frac1 = c.use_lam ? fmin(c.interp_fac * lam, 1.0) : fmin(c.interp_fac * nneigh * c.i_n_full, 1.0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double acc_drop = 0.5 * Vd[k] * Vn / sos2 * c.gamma_g * cellP * (kPi * c.L * c.L * 0.25) * Cdi / mass;
        const double acc_kern = -Pi * c.A_plate * nn[k] / mass;
        const double acc_skin = 0.5 * Vpn * Vpar[k] / sos2 * c.gamma_g * cellP * Cf * c.A_plate / mass;
        out[k] = frac1 * (acc_kern + acc_skin) + (1.0 - frac1) * acc_drop;
    }
}

/* The skin-friction case of CalcAeroAcc (src/Aero.h:223-256): stagnation
 * pressure of the normal velocity and seventh-power-law friction of the
 * tangential one on the plate area, blended with the droplet drag by
 * min(1.5 lambda, 1); nothing unless the gas moves into the surface. */
template <int D>
__device__ void skin_friction(const Consts& c, const double* Vd, const double* nrm, double lam, double mass,
                              double* out)
{
    double nn[D];
    normalized_d<D>(nrm, nn);
    double Vnorm = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) Vnorm += Vd[k] * nn[k];
    if (!(Vnorm > 0.001)) {
#pragma unroll
        for (int k = 0; k < D; ++k) out[k] = 0.0;
        return;
    }
    const double Vn = sqrt(dotd<D>(Vd, Vd));
    const double Re = c.rho_g * Vn * c.L / c.mu_g;
    double Vpar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) Vpar[k] = Vd[k] - fabs(Vnorm) * nn[k];
    const double Cf = 0.027 / pow(Re, 1.0 / 7.0);
    const double frac2 = fmin(1.5 * lam, 1.0);
    const double frac1 = 1.0 - frac2;
    const double Cdi = sphere_cd(Re);
    const double Vpn = sqrt(dotd<D>(Vpar, Vpar));
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double acc_press = 0.5 * c.rho_g * Vnorm * Vnorm * c.A_plate * nn[k] / mass;
        const double acc_skin = 0.5 * c.rho_g * Vpn * Cf * c.A_plate * Vpar[k] / mass;
        const double acc_drop = 0.5 * c.rho_g * Vn * Vd[k] * (kPi * c.L * c.L / 4) * Cdi / mass;
        out[k] = frac2 * (acc_press + acc_skin) + frac1 * acc_drop;
    }
}


/* Occupancy request for the force passes: 5 waves per SIMD (at most 102
 * VGPRs) measured faster than the compiler's 4 at 97 VGPRs. */
#ifndef FJ_FORCE_OCC
#define FJ_FORCE_OCC __attribute__((amdgpu_waves_per_eu(5, 8)))
#endif

/* LDS tile of the FP64 force pass (parity path), one array per value: x, m,
 * v, p/rho^2, xb (FR), vPert, rho; 1/rho and m/rho are recomputed per pair
 * with the same operations as launch_aux, so they are bit-identical to the
 * stored ones. */
template <bool FR>
struct ForceTile {
    double P[8][TILE];
    double XB[3][FR ? TILE : 1];
    double EA[4][TILE];
};

/* Pair sums of get_acc_and_Rrho_on_i (src/Resid.cpp:243-418) for fluid
 * slots, in FP64: writes the pair sum of the acceleration to acc (the
 * aerodynamic, artificial-viscosity and gravity terms are added by
 * k_nb_update) and the full Rrho.  The pressure, viscous and
 * surface-tension terms of a pair are combined before accumulation:
 * acc_i = sum_j (visc + ST / m_i - m_j (P_i + P_j) g).
 * FR: the positions have moved since the build, r comes from Aux::xb. */
template <int D, bool FR>
__global__ __launch_bounds__(TPB) FJ_FORCE_OCC void k_force_t(Consts c, DevState s, Aux ax,
                                                              const int32_t* __restrict__ part_id, TileGrid tg,
                                                              NbrList nl, const DevScalars* sc)
{
    __shared__ ForceTile<FR> ft;
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    double npd = sc->npd;
    if (tg.lgrp) {
        /* independent simulations: a block holds one group's targets; a
         * group that has converged keeps its rates */
        const int32_t g = block_group<D>(tg, h.b);
        if (g < 0 || !tg.gactive[g]) return;
        npd = tg.gnpd[g];
    }
    const double ste = -(0.5 * c.sig / c.lam_st) / (npd * npd);
    const double pi3o4H = 3.0 * kPi / 4.0 / c.H;
    const double inv_H = 1.0 / c.H, Ht = c.H, Wc = c.Wc;
    const double eps2 = 0.001 * c.H_sq;
    const double nu = c.nu;
    const double contang_fac = 1.0 + 0.5 * cos(0.5 * kPi * 7.0 / 9.0);
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        double xi[D], vi[D], xbi[D], vpi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xi[a] = active ? s.x[i * D + a] : 0.0;
            vi[a] = active ? s.v[i * D + a] : 0.0;
            xbi[a] = (FR && active) ? ax.xb[i * D + a] : 0.0;
            vpi[a] = (fluid && c.ale) ? s.vPert[i * D + a] : 0.0;
        }
        const double rhoi = fluid ? s.rho[i] : 1.0, mi = fluid ? s.m[i] : 1.0;
        const int32_t bi = fluid ? s.b[i] : 0;
        const double Pi = fluid ? ax.P[i] : 0.0;
        const double iri = fluid ? ax.invrho[i] : 0.0;
        const double inv_mi = 1.0 / mi;
        const bool st_on = (!c.ale) || (fluid && s.surfzone[i] == 1);
        const bool bound_i = bi == FJSPH_BOUND;
        double acc[D];
#pragma unroll
        for (int a = 0; a < D; ++a) acc[a] = 0.0;
        double Rrho_ = 0.0, Rrhoc_ = 0.0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    ft.P[a][t] = s.x[(int64_t)sl * D + a];
                    ft.P[4 + a][t] = s.v[(int64_t)sl * D + a];
                    if constexpr (FR) ft.XB[FR ? a : 0][FR ? t : 0] = ax.xb[(int64_t)sl * D + a];
                    ft.EA[a][t] = c.ale ? s.vPert[(int64_t)sl * D + a] : 0.0;
                }
                ft.P[3][t] = s.b[sl] == FJSPH_BOUND ? -s.m[sl] : s.m[sl];
                ft.P[7][t] = ax.P[sl];
                ft.EA[3][t] = s.rho[sl];
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                double Rji[D], Vji[D], vjr[D], vpj[D], xj[D], d2;
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    xj[a] = ft.P[a][t];
                    Rji[a] = xj[a] - xi[a];
                    Vji[a] = ft.P[4 + a][t] - vi[a];
                    vjr[a] = ft.P[4 + a][t];
                    vpj[a] = ft.EA[a][t];
                }
                if constexpr (FR) {
                    double xbj[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) xbj[a] = ft.XB[FR ? a : 0][FR ? t : 0];
                    d2 = exact_dist2<D>(xbi, xbj);
                } else {
                    d2 = exact_dist2<D>(xi, xj);
                }
                const double rr = d2, r = sqrt(d2);
                const double msg = ft.P[3][t];
                const double Pj = ft.P[7][t];
                const double rhoj = ft.EA[3][t];
                const double irj = 1.0 / rhoj;
                const double volj = fabs(msg) / rhoj;
                const bool bound_j = msg < 0.0;
                const double mj = bound_j ? -msg : msg;
                const double idist2 = 1.0 / (rr + eps2);
                double W, fg;
                wendland<double>(r, Ht, inv_H, Wc, W, fg);
                double g[D];
#pragma unroll
                for (int a = 0; a < D; ++a) g[a] = fg * Rji[a];
                const double pc = mj * (Pi + Pj);
                double rdotg = Rji[0] * g[0];
#pragma unroll
                for (int a = 1; a < D; ++a) rdotg += Rji[a] * g[a];
                const double vf = mj * nu * (iri + irj) * rdotg * idist2;
                double sf = 0.0;
                if (st_on) {
                    const double fac = (bound_i || bound_j) ? contang_fac : 1.0;
                    sf = ste * fac * cos(pi3o4H * r) * (1.0 / r) * inv_mi;
                }
#pragma unroll
                for (int a = 0; a < D; ++a) acc[a] += vf * Vji[a] + sf * Rji[a] - pc * g[a];
                if (c.ale) {
                    double aj = 0.0, ai = 0.0, cd = 0.0, cont = 0.0, c2 = 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) {
                        aj += vpj[a] * g[a];
                        ai += vpi[a] * g[a];
                        cd += (vpj[a] - vpi[a]) * g[a];
                        cont += (Vji[a] + (vpj[a] - vpi[a])) * g[a];
                        c2 += (rhoj * vpj[a] + rhoi * vpi[a]) * g[a];
                    }
#pragma unroll
                    for (int a = 0; a < D; ++a) acc[a] += (vjr[a] * aj + vi[a] * ai - vi[a] * cd) * volj;
                    Rrho_ -= cont * volj;
                    Rrhoc_ += c2 * volj;
                } else {
                    double vg = Vji[0] * g[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) vg += Vji[a] * g[a];
                    Rrho_ -= volj * vg;
                }
            FJ_TILE_ENTRIES_END(cu, nl)
        }
        if (!fluid) continue;
#pragma unroll
        for (int a = 0; a < D; ++a) s.acc[i * D + a] = acc[a];
        s.Rrho[i] = c.ale ? (Rrho_ * rhoi + Rrhoc_ + s.deltaD[i]) : (Rrho_ * rhoi + s.deltaD[i]);
    }
}

/* LDS tile of the mixed-precision force pass, one record per union slot,
 * four float4 reads per pair (three without ALE):
 *   A {x - xref, m (negative for BOUND)}, B {v - vref, p/rho^2},
 *   X {xb - xbref (build positions, FR only), 1/rho}, E {vPert, rho};
 *   V_j = |m_j| / rho_j is formed as |m_j| * (1/rho_j). */
template <bool ALE>
struct ForceTileM {
    float4 A[TILE], B[TILE], X[TILE], E[ALE ? TILE : 1];
};

/* Mixed-precision form of k_force_t: pair terms in FP32 on coordinates and
 * velocities relative to the block's first target (split_rel), each sum
 * carried in FP32 over one LDS tile (a few dozen pairs) and the tile sums
 * added into FP64 totals.
 *
 * The ALE terms are rearranged (exact in real arithmetic): with a_j = vPert_j
 * . g, a_i = vPert_i . g and V_ji = v_j - v_i, the reference's
 * (v_j a_j + v_i a_i - v_i (a_j - a_i)) V_j equals a_j V_j V_ji + 2 a_i V_j
 * v_i, and its continuity terms -(V_ji . g + a_j - a_i) V_j and (rho_j a_j +
 * rho_i a_i) V_j hold a_i only through sum_j V_j g.  So the pair loop carries
 * a_j and S = sum_j V_j g, and vPert_i . S enters the three sums once per
 * target. */
template <int D, bool FR, bool ALE>
__global__ __launch_bounds__(TPB) FJ_FORCE_OCC void k_force_m(Consts c, DevState s, Aux ax,
                                                              const int32_t* __restrict__ part_id, TileGrid tg,
                                                              NbrList nl, const DevScalars* sc)
{
    __shared__ ForceTileM<ALE> ft;
    __shared__ BlockHead h;
    block_head<D>(tg, h);
    const int U = h.pre[h.nseg];
    const int NT = n_tiles(U);
    /* reference point of the block (its first target) for the FP32
     * relative coordinates */
    double xref[D], vref[D], xbref[D];
    {
        const int64_t i0 = tg.order[h.p0 < h.p1 ? h.p0 : 0];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            xref[a] = s.x[i0 * D + a];
            vref[a] = s.v[i0 * D + a];
            xbref[a] = FR ? ax.xb[i0 * D + a] : 0.0;
        }
    }
    double npd = sc->npd;
    if (tg.lgrp) {
        const int32_t g = block_group<D>(tg, h.b);
        if (g < 0 || !tg.gactive[g]) return;
        npd = tg.gnpd[g];
    }
    const float ste = float(-(0.5 * c.sig / c.lam_st) / (npd * npd));
    const float pi3o4H = float(3.0 * kPi / 4.0 / c.H);
    const float inv_H = float(1.0 / c.H);
    const float fgc = float(5.0 * c.Wc / c.H_sq);
    /* ALE form of W'(r)/r = 5 Wc / H^2 (1 - r / 2H)^3: (kc - kc r / 2H)^3,
     * kc the cube root of 5 Wc / H^2 */
    const double kcd = cbrt(5.0 * c.Wc / c.H_sq);
    const float kc = float(kcd), kq = float(-0.5 * kcd / c.H), rmin = float(1e-12 * c.H);
    const float eps2 = float(0.001 * c.H_sq);
    const float nu = float(c.nu);
    const float contang_fac = float(1.0 + 0.5 * cos(0.5 * kPi * 7.0 / 9.0));
    for (int pb = h.p0; pb < h.p1; pb += TPB) {
        const int p = pb + threadIdx.x;
        const bool active = p < h.p1;
        const int64_t i = active ? tg.vorder[p] : 0;
        const bool fluid = active && part_id[i] >= c.n_bound;
        float xit[D], xlo[D], vit[D], vlo[D], xbit[D], xblo[D], vpi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            split_rel(active ? s.x[i * D + a] : xref[a], xref[a], xit[a], xlo[a]);
            split_rel(active ? s.v[i * D + a] : vref[a], vref[a], vit[a], vlo[a]);
            if constexpr (FR) split_rel(active ? ax.xb[i * D + a] : xbref[a], xbref[a], xbit[a], xblo[a]);
            else xbit[a] = xblo[a] = 0.f;
            vpi[a] = (ALE && fluid) ? float(s.vPert[i * D + a]) : 0.f;
        }
        const double rhoi = fluid ? s.rho[i] : 1.0, mi = fluid ? s.m[i] : 1.0;
        const int32_t bi = fluid ? s.b[i] : 0;
        const float Pi = fluid ? float(ax.P[i]) : 0.f;
        const float iri = fluid ? float(ax.invrho[i]) : 0.f;
        const float stm = ste * float(1.0 / mi);
        /* surface tension: everywhere without ALE, in the surface zone with
         * it; no terms at all when sigma = 0 */
        const bool st_on = c.sig != 0.0 && (!ALE || (fluid && s.surfzone[i] == 1));
        const bool bound_i = bi == FJSPH_BOUND;
        double acc[D];
#pragma unroll
        for (int a = 0; a < D; ++a) acc[a] = 0.0;
        double Rrho_ = 0.0, Rrhoc_ = 0.0, vS = 0.0;
        Rows cu;
        cu.init(nl, p, active ? nl.count[i] : 0, fluid);
        TileSlots ts;
        ts.load(tg, h, 0, NT, U);
        for (int tk = 0, base = 0; tk < NT; ++tk, base += TILE) {
            const int nt = tile_len(tk, U, NT);
            __syncthreads();
            for (int t = threadIdx.x; t < nt; t += TPB) {
                const int32_t sl = ts.at(t);
                if (sl < 0) continue;
                float pa[3] = {0.f, 0.f, 0.f}, va[3] = {0.f, 0.f, 0.f}, xa[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    pa[a] = float(s.x[(int64_t)sl * D + a] - xref[a]);
                    va[a] = float(s.v[(int64_t)sl * D + a] - vref[a]);
                    if constexpr (FR) xa[a] = float(ax.xb[(int64_t)sl * D + a] - xbref[a]);
                }
                const float4 q = ax.fq[sl];
                ft.A[t] = make_float4(pa[0], pa[1], pa[2], q.x);
                ft.B[t] = make_float4(va[0], va[1], va[2], q.y);
                ft.X[t] = make_float4(xa[0], xa[1], xa[2], q.z);
                if constexpr (ALE) ft.E[t] = ax.fale[sl];
            }
            ts.load(tg, h, tk + 1, NT, U);
            __syncthreads();
            float fa[3] = {0.f, 0.f, 0.f}, fS[3] = {0.f, 0.f, 0.f}, fr = 0.f, frc = 0.f;
            FJ_TILE_ENTRIES_BEGIN(cu, nl, base)
                const float4 A = ft.A[t], B = ft.B[t], X = ft.X[t];
                const float pj[3] = {A.x, A.y, A.z}, vj[3] = {B.x, B.y, B.z}, xbj[3] = {X.x, X.y, X.z};
                float Rji[D], Vji[D];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    Rji[a] = (pj[a] - xit[a]) - xlo[a];
                    Vji[a] = (vj[a] - vit[a]) - vlo[a];
                }
                float rr;
                if constexpr (FR) {
                    float Rb[D];
#pragma unroll
                    for (int a = 0; a < D; ++a) Rb[a] = (xbj[a] - xbit[a]) - xblo[a];
                    rr = Rb[0] * Rb[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rb[a] * Rb[a];
                } else {
                    rr = Rji[0] * Rji[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) rr += Rji[a] * Rji[a];
                }
                const float r = fast_sqrt<float>(rr);
                const float msg = A.w, Pj = B.w, irj = X.w;
                const bool bound_j = msg < 0.f;
                const float mj = fabsf(msg);
                const float volj = mj * irj;
                const float idist2 = fast_rcp<float>(rr + eps2);
                /* W'(r)/r is zero below r = 1e-12 H (src/Kernel.h:48-61); with r
                 * from the current positions such an FP32 distance is zero and so
                 * is R_ji, which carries every term of the pair */
                float fg;
                if constexpr (ALE) {
                    const float tq = fmaf(r, kq, kc);
                    fg = (FR && r < rmin) ? 0.f : tq * tq * tq;
                } else {
                    const float qh = r * inv_H;
                    const float tq = 1.f - 0.5f * qh;
                    fg = (FR && qh < 1e-12f) ? 0.f : fgc * (tq * tq * tq);
                }
                float g[D];
#pragma unroll
                for (int a = 0; a < D; ++a) g[a] = fg * Rji[a];
                const float pc = mj * (Pi + Pj);
                /* with ALE and r from the current positions, R_ji . g is
                 * W'(r)/r r^2 */
                constexpr bool rdot_rr = ALE && !FR;
                float rdotg = rdot_rr ? fg * rr : Rji[0] * g[0], vg = Vji[0] * g[0];
#pragma unroll
                for (int a = 1; a < D; ++a) {
                    if constexpr (!rdot_rr) rdotg += Rji[a] * g[a];
                    vg += Vji[a] * g[a];
                }
                float vf = mj * nu * (iri + irj) * rdotg * idist2;
                /* surface tension: with ALE only the surface zone has it, so
                 * its terms go to the sums inside the branch, which most waves
                 * skip; without ALE every fluid target has it and the terms
                 * join the momentum sum below */
                float sf = 0.f;
                if (st_on) {
                    const float fac = (bound_i || bound_j) ? contang_fac : 1.f;
                    sf = stm * fac * fast_cos<float>(pi3o4H * r) * fast_rcp<float>(r);
                    if constexpr (ALE) {
#pragma unroll
                        for (int a = 0; a < D; ++a) fa[a] += sf * Rji[a];
                    }
                }
                if constexpr (ALE) {
                    const float4 E = ft.E[t];
                    const float vpj[3] = {E.x, E.y, E.z};
                    float aj = vpj[0] * g[0];
#pragma unroll
                    for (int a = 1; a < D; ++a) aj += vpj[a] * g[a];
                    vf += volj * aj;
                    fr += volj * (vg + aj);
                    frc += (volj * E.w) * aj;
#pragma unroll
                    for (int a = 0; a < D; ++a) fS[a] += volj * g[a];
                } else {
                    fr += volj * vg;
                }
                if constexpr (ALE) {
#pragma unroll
                    for (int a = 0; a < D; ++a) fa[a] += vf * Vji[a] - pc * g[a];
                } else {
#pragma unroll
                    for (int a = 0; a < D; ++a) fa[a] += vf * Vji[a] + sf * Rji[a] - pc * g[a];
                }
            FJ_TILE_ENTRIES_END(cu, nl)
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] += double(fa[a]);
            Rrho_ -= double(fr);
            if constexpr (ALE) {
                float vs = vpi[0] * fS[0];
#pragma unroll
                for (int a = 1; a < D; ++a) vs += vpi[a] * fS[a];
                Rrhoc_ += double(frc);
                vS += double(vs);
            }
        }
        if (!fluid) continue;
        if constexpr (ALE) {
#pragma unroll
            for (int a = 0; a < D; ++a) acc[a] += 2.0 * s.v[i * D + a] * vS;
            Rrho_ += vS;
            Rrhoc_ += rhoi * vS;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) s.acc[i * D + a] = acc[a];
        s.Rrho[i] = ALE ? (Rrho_ * rhoi + Rrhoc_ + s.deltaD[i]) : (Rrho_ * rhoi + s.deltaD[i]);
    }
}

/* Per-slot tail of get_acc_and_Rrho_on_i (aerodynamic force, artificial
 * viscosity and gravity added to the pair sum, src/Resid.cpp:381-418) and
 * the Newmark-Beta update of Do_NB_Iter (src/Newmark_Beta.cpp:137-240) with
 * the per-block partial of the squared position change. */
template <int D>
__global__ __launch_bounds__(TPB) void k_nb_update(Consts c, DevState s, DevState pn, NextState nx, int do_update,
                                                   const int32_t* __restrict__ part_id,
                                                   const int32_t* __restrict__ blk, DevBlocks bl,
                                                   const int32_t* __restrict__ near_inlet,
                                                   const int32_t* __restrict__ nnb, int64_t n, double dt,
                                                   double* partial, DevGroups gr)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    double resid = 0.0;
    bool held = false; /* a group that has converged keeps its state */
    if (gr.grp && i < n) {
        const int32_t g = gr.grp[i];
        held = !gr.active[g];
        dt = gr.dt[g];
    }
    if (held) {
        if (do_update) {
#pragma unroll
            for (int a = 0; a < D; ++a) { nx.x[i * D + a] = s.x[i * D + a]; nx.v[i * D + a] = s.v[i * D + a]; }
            nx.rho[i] = s.rho[i];
            nx.p[i] = s.p[i];
            gr.d2[i] = 0.0;
        }
    } else if (i < n) {
        const bool fluid = part_id[i] >= c.n_bound;
        double xi[D], vi[D], acc_new[D];
        load_vec<D>(s.x, i, xi);
        load_vec<D>(s.v, i, vi);
        if (fluid) {
            double acc[D];
            load_vec<D>(s.acc, i, acc);
            if (s.cellID[i] != -1) {
                double Vd[D], aero[D];
#pragma unroll
                for (int a = 0; a < D; ++a) Vd[a] = s.cellV[i * D + a] - vi[a];
                if (c.aero_case == FJSPH_GISSLER) {
                    gissler<D>(c, Vd, s.cellRho[i], s.cellP[i], s.m[i], s.lam_nb[i], double(nnb[i] + 1), s.woccl[i],
                               aero);
                } else if (c.aero_case == FJSPH_INDUCED_PRESSURE) {
                    double nrm[D];
                    load_vec<D>(s.norm, i, nrm);
                    induced_pressure<D>(c, Vd, nrm, s.lam_nb[i], double(nnb[i] + 1), s.norm_curve[i], s.cellP[i],
                                        s.cellRho[i], s.m[i], aero);
                } else if (c.aero_case == FJSPH_SKIN_FRIC) {
                    double nrm[D];
                    load_vec<D>(s.norm, i, nrm);
                    skin_friction<D>(c, Vd, nrm, s.lam_nb[i], s.m[i], aero);
                } else {
#pragma unroll
                    for (int a = 0; a < D; ++a) aero[a] = 0.0;
                }
#pragma unroll
                for (int a = 0; a < D; ++a) { acc[a] += aero[a]; s.Af[i * D + a] = aero[a]; }
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                acc_new[a] = acc[a] + s.aVisc[i * D + a] + c.grav[a];
                s.acc[i * D + a] = acc_new[a];
            }
        }
        if (do_update) {
            /* Newmark-Beta update (beta = 1/4, gamma = 1/2) */
            const double g1 = 0.5, g2 = 0.5, b1 = 0.25, b2 = 0.25;
            const double dt2 = dt * dt;
            double xo[D], vo[D], rho_o = s.rho[i], p_o = s.p[i];
#pragma unroll
            for (int a = 0; a < D; ++a) { xo[a] = xi[a]; vo[a] = vi[a]; }
            const int32_t bcur = s.b[i];
            if (fluid) {
                if (bcur > FJSPH_BUFFER && bcur != FJSPH_OUTLET) {
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
                    rho_o = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dt * (g1 * s.Rrho[i] + g2 * pn.Rrho[i])));
                    p_o = eos_pressure(c, rho_o);
                } else if (bcur == FJSPH_OUTLET) {
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
            if (gr.grp) gr.d2[i] = resid;
        }
    }
    if (do_update && !gr.grp) {
        const double bs = block_reduce(resid, OpSum(), lds);
        if (threadIdx.x == 0) partial[blockIdx.x] = bs;
    }
}

/* ------------------------------------------------------- Runge-Kutta 4 */

/* Fluid update of do_runge_kutta_intermediate_step (final = 0) and
 * do_runge_kutta_final_step (final = 1) (src/Runge_Kutta.cpp:28-457) on the
 * stage state s, whose acc and Rrho were just computed; the x update reads
 * the velocity s held before this stage.  The final step combines the
 * previous step's acceleration (pn.acc) with stages 1-3, as the reference
 * does. */
template <int D>
__global__ __launch_bounds__(TPB) void k_rk_update(Consts c, DevState s, DevState pn, DevState s1, DevState s2,
                                                   DevState s3, const int32_t* __restrict__ part_id, int64_t n,
                                                   double dt, int final_stage)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    const int32_t b = s.b[i];
    if (!final_stage) {
        if (b <= FJSPH_BUFFER) return;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int64_t k = i * D + a;
            const double vel = c.ale ? (s.v[k] + s.vPert[k]) : s.v[k];
            s.x[k] = pn.x[k] + dt * vel;
            s.v[k] = pn.v[k] + dt * s.acc[k];
        }
        const double rho = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dt * s.Rrho[i]));
        s.rho[i] = rho;
        s.p[i] = eos_pressure(c, rho);
        return;
    }
    const double h6 = dt / 6.0;
    if (b > FJSPH_BUFFER && b != FJSPH_OUTLET) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int64_t k = i * D + a;
            double vs;
            if (c.ale)
                vs = (((pn.v[k] + pn.vPert[k]) + 2.0 * (s1.v[k] + s1.vPert[k])) + 2.0 * (s2.v[k] + s2.vPert[k])) +
                     (s3.v[k] + s3.vPert[k]);
            else
                vs = ((pn.v[k] + 2.0 * s1.v[k]) + 2.0 * s2.v[k]) + s3.v[k];
            s.x[k] = pn.x[k] + h6 * vs;
            s.v[k] = pn.v[k] + h6 * (((pn.acc[k] + 2.0 * s1.acc[k]) + 2.0 * s2.acc[k]) + s3.acc[k]);
        }
        const double dr = ((pn.Rrho[i] + 2.0 * s1.Rrho[i]) + 2.0 * s2.Rrho[i]) + s3.Rrho[i];
        const double rho = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + h6 * dr));
        s.rho[i] = rho;
        s.p[i] = eos_pressure(c, rho);
    } else if (b == FJSPH_OUTLET) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const int64_t k = i * D + a;
            s.x[k] = pn.x[k] + dt * s.v[k];
        }
    }
}

/* Density of DBC and ghost boundary particles in a Runge-Kutta stage (the
 * solver part of rk_boundary, src/Runge_Kutta.cpp:69-160 and 281-380). */
template <int D>
__global__ __launch_bounds__(TPB) void k_rk_bnd_rho(Consts c, DevState s, DevState pn, DevState s1, DevState s2,
                                                    DevState s3, const int32_t* __restrict__ blk, DevBlocks bl,
                                                    const int32_t* __restrict__ near_inlet, int64_t n, double dt,
                                                    int final_stage)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    const int bk = blk[i];
    if (bk < 0) return;
    const int solver = bl.solver[bk];
    if (solver != FJSPH_DBC && solver != FJSPH_GHOST) return;
    const double drho = final_stage ? (((pn.Rrho[i] + 2.0 * s1.Rrho[i]) + 2.0 * s2.Rrho[i]) + s3.Rrho[i]) : s.Rrho[i];
    const double fac = final_stage ? (dt / 6.0) : dt;
    const bool near = solver == FJSPH_GHOST && near_inlet[i];
    const double lo = near ? c.rho0 : c.rho_min;
    const double rho = fmax(lo, fmin(c.rho_max, pn.rho[i] + fac * drho));
    s.rho[i] = rho;
    s.p[i] = eos_pressure(c, rho);
    if (solver == FJSPH_DBC && final_stage) s.Rrho[i] = s3.Rrho[i];
    if (near) s.Rrho[i] = fmax(0.0, s.Rrho[i]);
}

/* Per-block partials of sum |xb - xa|^2 over fluid slots (Check_RK_Error,
 * src/Runge_Kutta.cpp:10-26). */
template <int D>
__global__ __launch_bounds__(TPB) void k_sqdiff(const double* __restrict__ xa, const double* __restrict__ xb,
                                                const int32_t* __restrict__ part_id, int64_t n_bound, int64_t n,
                                                double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    double e = 0.0;
    if (i < n && part_id[i] >= n_bound) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const double d = xb[i * D + a] - xa[i * D + a];
            e += d * d;
        }
    }
    const double bs = block_reduce(e, OpSum(), lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

void launch_rk_update(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                      const int32_t* part_id, int64_t n, double dt, int final_stage, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_rk_update<2>, g, b, 0, st, c, s, pn, s1, s2, s3, part_id, n, dt, final_stage);
    else hipLaunchKernelGGL(k_rk_update<3>, g, b, 0, st, c, s, pn, s1, s2, s3, part_id, n, dt, final_stage);
}

void launch_rk_bnd_rho(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                       const int32_t* blk, DevBlocks bl, const int32_t* near_inlet, int64_t n, double dt,
                       int final_stage, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_rk_bnd_rho<2>, g, b, 0, st, c, s, pn, s1, s2, s3, blk, bl, near_inlet, n, dt, final_stage);
    else
        hipLaunchKernelGGL(k_rk_bnd_rho<3>, g, b, 0, st, c, s, pn, s1, s2, s3, blk, bl, near_inlet, n, dt, final_stage);
}

void launch_sqdiff(int dim, const double* xa, const double* xb, const int32_t* part_id, int64_t n_bound, int64_t n,
                   double* partial, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_sqdiff<2>, g, b, 0, st, xa, xb, part_id, n_bound, n, partial);
    else hipLaunchKernelGGL(k_sqdiff<3>, g, b, 0, st, xa, xb, part_id, n_bound, n, partial);
}

void launch_force(int dim, const Consts& c, DevState s, Aux a, const int32_t* part_id, TileGrid tg, NbrList nl,
                  int r_from_x, const DevScalars* sc, hipStream_t st)
{
    const dim3 g((unsigned)tg.nblocks), b(TPB);
#define FJ_FORCE_T(DD, FF) hipLaunchKernelGGL((k_force_t<DD, FF>), g, b, 0, st, c, s, a, part_id, tg, nl, sc)
#define FJ_FORCE_M(DD, FF, AA) hipLaunchKernelGGL((k_force_m<DD, FF, AA>), g, b, 0, st, c, s, a, part_id, tg, nl, sc)
#define FJ_FORCE_MA(DD, FF) \
    do { if (c.ale) FJ_FORCE_M(DD, FF, true); else FJ_FORCE_M(DD, FF, false); } while (0)
    if (dim == 2) {
        if (nl.f32) { if (r_from_x) FJ_FORCE_MA(2, false); else FJ_FORCE_MA(2, true); }
        else { if (r_from_x) FJ_FORCE_T(2, false); else FJ_FORCE_T(2, true); }
    } else {
        if (nl.f32) { if (r_from_x) FJ_FORCE_MA(3, false); else FJ_FORCE_MA(3, true); }
        else { if (r_from_x) FJ_FORCE_T(3, false); else FJ_FORCE_T(3, true); }
    }
#undef FJ_FORCE_MA
#undef FJ_FORCE_M
#undef FJ_FORCE_T
}

void launch_nb_update(int dim, const Consts& c, DevState s, DevState pn, double* next_x, double* next_v,
                      double* next_rho, double* next_p, const int32_t* part_id, const int32_t* blk, DevBlocks bl,
                      const int32_t* near_inlet, NbrList nl, int64_t n, double dt, double* partial, hipStream_t st,
                      DevGroups gr)
{
    NextState nx{next_x, next_v, next_rho, next_p};
    const int upd = next_x != nullptr;
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_nb_update<2>, g, b, 0, st, c, s, pn, nx, upd, part_id, blk, bl, near_inlet, nl.nnb, n, dt,
                           partial, gr);
    else
        hipLaunchKernelGGL(k_nb_update<3>, g, b, 0, st, c, s, pn, nx, upd, part_id, blk, bl, near_inlet, nl.nnb, n, dt,
                           partial, gr);
}

}  // namespace fj
