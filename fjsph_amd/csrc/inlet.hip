/*
 * Inlets, deletion and the pipe exit (scope table "next" row 1):
 *
 *   k_pipe_outlet   Check_Pipe_Outlet (src/Containment.cpp:822-890): a PIPE
 *                   particle past its block's aerodynamic plane becomes FREE
 *   k_inlet_buffer  the inlet-block buffer update of Do_NB_Iter
 *                   (src/Newmark_Beta.cpp:243-292): dynamic inlets place the
 *                   buffer particles of a column behind its back particle
 *                   and give them its v, rho, p; fixed-velocity inlets move
 *                   them with their velocity and integrate their density
 *   k_inlet_back,   update_buffer_region (src/shapes/inlet.cpp:578-640): a
 *   k_inlet_shift,  back particle past the insertion plane becomes PIPE, the
 *   k_inlet_insert  column's buffer moves up one place and a new BUFFER
 *                   particle is made behind the last one from its state
 *   k_delete_flag   the deletion plane of update_data
 *                   (src/Integration.cpp:126-205)
 *
 * Inlet columns are kept per particle, not as index lists: col (the
 * column, -1 outside inlets) and depth (0 for the back particle, 1..nbuf
 * for the buffer behind it) travel with the particle through every
 * permutation.  A plane is (n, c): a particle crosses it when x . n > c.
 */
#include "device_common.hpp"

namespace fj {

template <int D>
__device__ __forceinline__ double dot_plane(const double* __restrict__ x, int64_t i, const double* pl)
{
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) s += x[i * D + a] * pl[a];
    return s;
}

template <int D>
__global__ __launch_bounds__(TPB) void k_pipe_outlet(DevState s, const int32_t* __restrict__ part_id,
                                                     const int32_t* __restrict__ blk, DevPlanes pl, int64_t n,
                                                     int64_t n_bound, int32_t* __restrict__ fresh, double lam_cutoff,
                                                     DevScalars* sc)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < n_bound || s.b[i] != FJSPH_PIPE) return;
    const int k = blk[i];
    if (k < 0) return;
    if (dot_plane<D>(s.x, i, pl.aero + 4 * k) > pl.aero[4 * k + 3]) {
        s.b[i] = FJSPH_FREE;
        /* with the mesh source, FirstCell for the new aero particles */
        if (fresh && s.lam_nb[i] < lam_cutoff) fresh[atomicAdd(&sc->fresh_n, 1)] = int32_t(i);
    }
}

/* Slot of every column's back particle (depth 0). */
__global__ __launch_bounds__(TPB) void k_inlet_backslot(const int32_t* __restrict__ col,
                                                        const int32_t* __restrict__ depth, int64_t n,
                                                        int32_t* __restrict__ back_slot)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || col[i] < 0 || depth[i] != 0) return;
    back_slot[col[i]] = int32_t(i);
}

/* After k_nb_update, in its grid: the next iterate of the buffer
 * particles, and their squared position change added to the block's
 * residual partial. */
template <int D>
__global__ __launch_bounds__(TPB) void k_inlet_buffer(Consts c, DevState s, DevState pn, NextState nx,
                                                      const int32_t* __restrict__ col,
                                                      const int32_t* __restrict__ depth,
                                                      const int32_t* __restrict__ blk,
                                                      const int32_t* __restrict__ back_slot, DevPlanes pl,
                                                      int64_t n, double dt, double dx, double* partial)
{
    __shared__ double lds[NWAVE];
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    double resid = 0.0;
    if (i < n && col[i] >= 0 && depth[i] > 0 && s.b[i] == FJSPH_BUFFER) {
        const int k = blk[i];
        double xo[D];
        if (pl.mode[k] == 1) {
            /* dynamic: off the column's back particle (its new iterate) */
            const int64_t bk = back_slot[col[i]];
            const double* nrm = pl.ins + 4 * k;
            double len = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) len += nrm[a] * nrm[a];
            len = sqrt(len);
#pragma unroll
            for (int a = 0; a < D; ++a) {
                xo[a] = nx.x[bk * D + a] - dx * double(depth[i]) * (nrm[a] / len);
                nx.v[i * D + a] = nx.v[bk * D + a];
            }
            nx.rho[i] = nx.rho[bk];
            nx.p[i] = nx.p[bk];
        } else {
            /* fixed velocity */
            const double rho = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dt * (0.5 * s.Rrho[i] + 0.5 * pn.Rrho[i])));
            nx.rho[i] = rho;
            nx.p[i] = eos_pressure(c, rho);
#pragma unroll
            for (int a = 0; a < D; ++a) xo[a] = pn.x[i * D + a] + dt * pn.v[i * D + a];
        }
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const double d = xo[a] - s.x[i * D + a];
            resid += d * d;
            nx.x[i * D + a] = xo[a];
        }
    }
    const double bs = block_reduce(resid, OpSum(), lds);
    if (threadIdx.x == 0 && bs != 0.0) partial[blockIdx.x] += bs;
}

/* The inlet buffer of a Runge-Kutta stage (src/Runge_Kutta.cpp:171-229,
 * 389-447), after k_rk_update: with fixed_vel_or_dynamic = 1 off the
 * column's back particle of this stage; otherwise the density from pn and
 * the particle's own stage rate (dt Rrho, or dt/6 of the four rates in the
 * last stage) and the position x_n + dt v_n.  The reference's branch for
 * the latter reads the buffer list of block `jj` (the buffer depth) and, in
 * the last stage, the rates of particle `ii` (the column number); the
 * block's own buffer and the particle's own rates are used here. */
template <int D>
__global__ __launch_bounds__(TPB) void k_rk_buffer(Consts c, DevState s, DevState pn, DevState s1, DevState s2,
                                                   DevState s3, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ depth,
                                                   const int32_t* __restrict__ blk,
                                                   const int32_t* __restrict__ back_slot, DevPlanes pl, int64_t n,
                                                   double dt, double dx, int final_stage)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || col[i] < 0 || depth[i] <= 0 || s.b[i] != FJSPH_BUFFER) return;
    const int k = blk[i];
    if (pl.mode[k] == 1) {
        const int64_t bk = back_slot[col[i]];
        const double* nrm = pl.ins + 4 * k;
        double len = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) len += nrm[a] * nrm[a];
        len = sqrt(len);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            s.x[i * D + a] = s.x[bk * D + a] - dx * double(depth[i]) * (nrm[a] / len);
            s.v[i * D + a] = s.v[bk * D + a];
        }
        s.rho[i] = s.rho[bk];
        s.p[i] = s.p[bk];
        return;
    }
    const double dr = final_stage ? (dt / 6.0) * (((pn.Rrho[i] + 2.0 * s1.Rrho[i]) + 2.0 * s2.Rrho[i]) + s3.Rrho[i])
                                  : dt * s.Rrho[i];
    const double rho = fmax(c.rho_min, fmin(c.rho_max, pn.rho[i] + dr));
    s.rho[i] = rho;
    s.p[i] = eos_pressure(c, rho);
#pragma unroll
    for (int a = 0; a < D; ++a) s.x[i * D + a] = pn.x[i * D + a] + dt * pn.v[i * D + a];
}

/* Back particles: crossed the insertion plane -> PIPE, column advances.
 * Also records the slot of each column's last buffer particle (depth
 * nbuf), the source of a new particle. */
template <int D>
__global__ __launch_bounds__(TPB) void k_inlet_back(DevState s, int32_t* __restrict__ col, int32_t* __restrict__ depth,
                                                    const int32_t* __restrict__ blk, DevPlanes pl, int64_t n,
                                                    int32_t* __restrict__ adv, int32_t* __restrict__ last_slot)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || col[i] < 0) return;
    const int k = blk[i];
    const int cc = col[i];
    if (depth[i] == pl.nbuf[k]) last_slot[cc] = int32_t(i);
    if (depth[i] != 0) return;
    const bool crossed = dot_plane<D>(s.x, i, pl.ins + 4 * k) > pl.ins[4 * k + 3];
    adv[cc] = crossed ? 1 : 0;
    if (crossed) {
        s.b[i] = FJSPH_PIPE;
        col[i] = -1;
        depth[i] = -1;
    }
}

/* Buffer particles of advancing columns move up one place; the first
 * becomes the column's back particle.  The last one (depth nbuf, nbuf > 1)
 * moves up only when a new particle replaces it (k_inlet_insert): the
 * reference shifts its buffer list left and leaves the last entry in
 * place, so a column that gets no new particle (max_points reached) lists
 * that particle twice, and the buffer pass places it at the deeper of the
 * two positions. */
__global__ __launch_bounds__(TPB) void k_inlet_shift(DevState s, int32_t* __restrict__ col, int32_t* __restrict__ depth,
                                                     const int32_t* __restrict__ blk, DevPlanes pl, int64_t n,
                                                     const int32_t* __restrict__ adv)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || col[i] < 0 || depth[i] <= 0 || !adv[col[i]]) return;
    const int32_t nb = pl.nbuf[blk[i]];
    if (depth[i] == nb && nb > 1) return;
    const int32_t d = depth[i] - 1;
    depth[i] = d;
    if (d == 0) s.b[i] = FJSPH_BACK;
}

/* New particles (one thread per column with a rank): slot n_old + rank,
 * made from the column's last buffer particle as the reference's
 * SPHPart(x, pj, BUFFER, id) does (src/Var.h:550-592: x, v, rho, p, m,
 * cellP, cellRho copied, no cell, everything else zero; the zeroing is
 * done before this kernel), placed dx behind it along the insertion
 * normal.  part_id is given to it afterwards. */
template <int D>
__global__ __launch_bounds__(TPB) void k_inlet_insert(DevState s, int32_t* __restrict__ blk,
                                                      int32_t* __restrict__ col, int32_t* __restrict__ depth,
                                                      const int32_t* __restrict__ col_block,
                                                      const int32_t* __restrict__ rank,
                                                      const int32_t* __restrict__ last_slot, DevPlanes pl,
                                                      int32_t ncols, int64_t n_old, double dx, int64_t id_base)
{
    const int cc = blockIdx.x * TPB + threadIdx.x;
    if (cc >= ncols || rank[cc] < 0) return;
    const int64_t src = last_slot[cc];
    const int64_t dst = n_old + rank[cc];
    const int k = col_block[cc];
    const double* nrm = pl.ins + 4 * k;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        s.x[dst * D + a] = s.x[src * D + a] - dx * nrm[a];
        s.v[dst * D + a] = s.v[src * D + a];
    }
    s.rho[dst] = s.rho[src];
    s.p[dst] = s.p[src];
    s.m[dst] = s.m[src];
    s.cellP[dst] = s.cellP[src];
    s.cellRho[dst] = s.cellRho[src];
    s.cellID[dst] = -3;
    s.b[dst] = FJSPH_BUFFER;
    s.pid[dst] = int32_t(id_base + rank[cc]);
    blk[dst] = k;
    col[dst] = cc;
    depth[dst] = pl.nbuf[k];
    depth[src] = pl.nbuf[k] - 1;
}

/* part_id of every particle at or past host index `from` moves up by
 * `delta` (the reference's vector insert at a block's end). */
__global__ __launch_bounds__(TPB) void k_shift_ids(int32_t* __restrict__ part_id, int64_t n, int32_t from, int32_t delta)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i < n && part_id[i] >= from) part_id[i] += delta;
}

/* New particles of block k get host ids first_new + (rank - rank_base). */
__global__ __launch_bounds__(TPB) void k_new_ids(int32_t* __restrict__ part_id, const int32_t* __restrict__ col_block,
                                                 const int32_t* __restrict__ rank, int32_t ncols, int32_t k,
                                                 int64_t n_old, int32_t rank_base, int32_t first_new)
{
    const int cc = blockIdx.x * TPB + threadIdx.x;
    if (cc >= ncols || rank[cc] < 0 || col_block[cc] != k) return;
    part_id[n_old + rank[cc]] = first_new + (rank[cc] - rank_base);
}

/* Deletion plane: flag[slot] = 1 and hflag[host id] = 1 for fluid
 * particles of blocks with an active plane that crossed it; per-block
 * counts. */
template <int D>
__global__ __launch_bounds__(TPB) void k_delete_flag(DevState s, const int32_t* __restrict__ part_id,
                                                     const int32_t* __restrict__ blk, DevPlanes pl, int64_t n,
                                                     int64_t n_bound, int32_t* __restrict__ keep,
                                                     int32_t* __restrict__ hflag, int32_t* __restrict__ blkcount)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    bool del = false;
    const int32_t h = part_id[i];
    if (h >= n_bound) {
        const int k = blk[i];
        if (k >= 0 && pl.del[4 * k + 3] != FJSPH_NO_PLANE && dot_plane<D>(s.x, i, pl.del + 4 * k) > pl.del[4 * k + 3])
            del = true;
        if (del) atomicAdd(blkcount + k, 1);
    }
    keep[i] = del ? 0 : 1;
    hflag[h] = del ? 1 : 0;
}

/* perm[new slot] = old slot for the kept slots (start = exclusive scan of
 * keep). */
__global__ __launch_bounds__(TPB) void k_compact_perm(const int32_t* __restrict__ keep,
                                                      const int32_t* __restrict__ start, int64_t n,
                                                      int32_t* __restrict__ perm)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i < n && keep[i]) perm[start[i]] = int32_t(i);
}

/* Host ids after deletion: minus the deleted ids below (hstart =
 * exclusive scan of hflag). */
__global__ __launch_bounds__(TPB) void k_compact_ids(int32_t* __restrict__ part_id, const int32_t* __restrict__ hstart,
                                                     int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i < n) part_id[i] -= hstart[part_id[i]];
}

/* ------------------------------------------------------------ launchers */

void launch_pipe_outlet(int dim, DevState s, const int32_t* part_id, const int32_t* blk, DevPlanes pl, int64_t n,
                        int64_t n_bound, hipStream_t st, int32_t* fresh, double lam_cutoff, DevScalars* sc)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_pipe_outlet<2>, g, b, 0, st, s, part_id, blk, pl, n, n_bound, fresh, lam_cutoff, sc);
    else
        hipLaunchKernelGGL(k_pipe_outlet<3>, g, b, 0, st, s, part_id, blk, pl, n, n_bound, fresh, lam_cutoff, sc);
}

void launch_inlet_backslot(const int32_t* col, const int32_t* depth, int64_t n, int32_t* back_slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_inlet_backslot, dim3(grid_of(n)), dim3(TPB), 0, st, col, depth, n, back_slot);
}

void launch_rk_buffer(int dim, const Consts& c, DevState s, DevState pn, DevState s1, DevState s2, DevState s3,
                      const int32_t* col, const int32_t* depth, const int32_t* blk, const int32_t* back_slot,
                      DevPlanes pl, int64_t n, double dt, double dx, int final_stage, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_rk_buffer<2>, g, b, 0, st, c, s, pn, s1, s2, s3, col, depth, blk, back_slot, pl, n, dt,
                           dx, final_stage);
    else
        hipLaunchKernelGGL(k_rk_buffer<3>, g, b, 0, st, c, s, pn, s1, s2, s3, col, depth, blk, back_slot, pl, n, dt,
                           dx, final_stage);
}

void launch_inlet_buffer(int dim, const Consts& c, DevState s, DevState pn, NextState nx, const int32_t* col,
                         const int32_t* depth, const int32_t* blk, const int32_t* back_slot, DevPlanes pl, int64_t n,
                         double dt, double dx, double* partial, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_inlet_buffer<2>, g, b, 0, st, c, s, pn, nx, col, depth, blk, back_slot, pl, n, dt, dx,
                           partial);
    else
        hipLaunchKernelGGL(k_inlet_buffer<3>, g, b, 0, st, c, s, pn, nx, col, depth, blk, back_slot, pl, n, dt, dx,
                           partial);
}

void launch_inlet_back(int dim, DevState s, int32_t* col, int32_t* depth, const int32_t* blk, DevPlanes pl, int64_t n,
                       int32_t* adv, int32_t* last_slot, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_inlet_back<2>, g, b, 0, st, s, col, depth, blk, pl, n, adv, last_slot);
    else hipLaunchKernelGGL(k_inlet_back<3>, g, b, 0, st, s, col, depth, blk, pl, n, adv, last_slot);
}

void launch_inlet_shift(DevState s, int32_t* col, int32_t* depth, const int32_t* blk, DevPlanes pl, int64_t n,
                        const int32_t* adv, hipStream_t st)
{
    hipLaunchKernelGGL(k_inlet_shift, dim3(grid_of(n)), dim3(TPB), 0, st, s, col, depth, blk, pl, n, adv);
}

void launch_inlet_insert(int dim, DevState s, int32_t* blk, int32_t* col, int32_t* depth, const int32_t* col_block,
                         const int32_t* rank, const int32_t* last_slot, DevPlanes pl, int32_t ncols, int64_t n_old,
                         double dx, int64_t id_base, hipStream_t st)
{
    const dim3 g(grid_of(ncols)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_inlet_insert<2>, g, b, 0, st, s, blk, col, depth, col_block, rank, last_slot, pl, ncols,
                           n_old, dx, id_base);
    else
        hipLaunchKernelGGL(k_inlet_insert<3>, g, b, 0, st, s, blk, col, depth, col_block, rank, last_slot, pl, ncols,
                           n_old, dx, id_base);
}

void launch_shift_ids(int32_t* part_id, int64_t n, int32_t from, int32_t delta, hipStream_t st)
{
    hipLaunchKernelGGL(k_shift_ids, dim3(grid_of(n)), dim3(TPB), 0, st, part_id, n, from, delta);
}

void launch_new_ids(int32_t* part_id, const int32_t* col_block, const int32_t* rank, int32_t ncols, int32_t k,
                    int64_t n_old, int32_t rank_base, int32_t first_new, hipStream_t st)
{
    hipLaunchKernelGGL(k_new_ids, dim3(grid_of(ncols)), dim3(TPB), 0, st, part_id, col_block, rank, ncols, k, n_old,
                       rank_base, first_new);
}

void launch_delete_flag(int dim, DevState s, const int32_t* part_id, const int32_t* blk, DevPlanes pl, int64_t n,
                        int64_t n_bound, int32_t* keep, int32_t* hflag, int32_t* blkcount, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2)
        hipLaunchKernelGGL(k_delete_flag<2>, g, b, 0, st, s, part_id, blk, pl, n, n_bound, keep, hflag, blkcount);
    else
        hipLaunchKernelGGL(k_delete_flag<3>, g, b, 0, st, s, part_id, blk, pl, n, n_bound, keep, hflag, blkcount);
}

void launch_compact_perm(const int32_t* keep, const int32_t* start, int64_t n, int32_t* perm, hipStream_t st)
{
    hipLaunchKernelGGL(k_compact_perm, dim3(grid_of(n)), dim3(TPB), 0, st, keep, start, n, perm);
}

void launch_compact_ids(int32_t* part_id, const int32_t* hstart, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_compact_ids, dim3(grid_of(n)), dim3(TPB), 0, st, part_id, hstart, n);
}

}  // namespace fj
