/*
 * Aerodynamic mesh lookup: FindCell / CheckCell (src/Containment.cpp:385-820)
 * with the face-crossing tests Crossings3D / Crossings2D
 * (src/Geometry.cpp:354-395, 485-575), one thread per particle.
 *
 * Candidate cells: instead of the reference's KD-tree over cell centres
 * (5, then 500 nearest), a uniform voxel grid lists every cell whose
 * bounding box overlaps the voxel, in ascending cell order.  A point inside
 * the mesh lies in the bounding box of its containing cell, so that cell is
 * always among its voxel's candidates; a point inside exactly one cell gets
 * the same cell as the reference.  The previous cell is tested first, as in
 * the reference.
 *
 * The signed volumes are evaluated with floating-point contraction off, in
 * the operation order of the CPU restatement, so their signs agree with it
 * bit for bit.
 */
#include "device_common.hpp"

namespace fj {

__device__ __forceinline__ const double* mvert(const DevMesh& M, int32_t v) { return M.verts + (int64_t)v * 3; }

__device__ int crossings3d(const DevMesh& M, int32_t f, const double* testp, const double* rayp)
{
    const int32_t* fv = M.face_v + M.face_off[f];
    const int32_t nv = M.face_off[f + 1] - M.face_off[f];
    const int flag1 = det4_rows(testp, mvert(M, fv[0]), mvert(M, fv[1]), mvert(M, fv[2])) < 0.0;
    const int flag2 = det4_rows(rayp, mvert(M, fv[0]), mvert(M, fv[1]), mvert(M, fv[2])) < 0.0;
    if (flag1 == flag2) return 0;
    const double* vtx0 = mvert(M, fv[nv - 1]);
    const double* vtx1 = mvert(M, fv[0]);
    const int flag3 = det4_rows(testp, vtx0, vtx1, rayp) < 0.0;
    for (int ii = 1; ii < 3; ++ii) {
        vtx0 = vtx1;
        vtx1 = mvert(M, fv[ii]);
        const int flag4 = det4_rows(testp, vtx0, vtx1, rayp) < 0.0;
        if (flag4 != flag3) return 0;
    }
    return 1;
}

__device__ int crossings2d(const DevMesh& M, int32_t f, const double* point)
{
#pragma clang fp contract(off)
    const int32_t* fv = M.face_v + M.face_off[f];
    const double tx = point[0], ty = point[1];
    const double* vtx0 = mvert(M, fv[0]);
    const double* vtx1 = mvert(M, fv[1]);
    const int yflag0 = vtx0[1] >= ty, yflag1 = vtx1[1] >= ty;
    if (yflag0 != yflag1)
        if (((vtx1[1] - ty) * (vtx1[0] - vtx0[0]) >= (vtx1[0] - tx) * (vtx1[1] - vtx0[1])) == yflag1) return 1;
    return 0;
}

template <int D>
__device__ int check_cell(const DevMesh& M, int32_t cell, const double* testp)
{
    const double rayp[3] = {testp[0] + 1e+5, testp[1], D == 3 ? testp[2] : 0.0};
    int line_flag = 0, inside_flag = 0;
    for (int32_t k = M.cell_off[cell]; k < M.cell_off[cell + 1]; ++k) {
        const int32_t f = M.cell_f[k];
        const int hit = D == 3 ? crossings3d(M, f, testp, rayp) : crossings2d(M, f, testp);
        if (hit) {
            inside_flag = !inside_flag;
            if (line_flag) break;
            line_flag = 1;
        }
    }
    return inside_flag;
}

/* Mesh vertices are stored with 3 components in both dimensions (z = 0 in
 * 2D), so the tests read them with one layout. */
template <int D>
__global__ __launch_bounds__(TPB) void k_find_cell(Consts c, DevState s, const int32_t* __restrict__ part_id,
                                                   DevMesh M, int64_t n, DevScalars* sc,
                                                   int32_t* __restrict__ lost_list)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n || part_id[i] < c.n_bound) return;
    if (s.b[i] != FJSPH_FREE || s.lam_nb[i] > c.lam_cutoff) {
        s.cellID[i] = -3;
        return;
    }
    double testp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) testp[a] = s.x[i * D + a];
    int32_t found = -1;
    const int32_t prev = s.cellID[i];
    /* still in its cell: ipt_n_failed is reset and internal kept; a cell
     * found by the search resets both */
    const bool same = prev >= 0 && prev < M.ncells && check_cell<D>(M, prev, testp);
    if (same) found = prev;
    if (!same) {
        int64_t vox = 0;
        bool inside = true;
#pragma unroll
        for (int a = D - 1; a >= 0; --a) {
            const double q = (testp[a] - M.glo[a]) * M.ginv;
            if (!(q >= 0.0) || q >= double(M.gdims[a])) inside = false;
            const int32_t g = inside ? int32_t(q) : 0;
            vox = vox * M.gdims[a] + g;
        }
        if (inside) {
            for (int32_t k = M.gstart[vox]; k < M.gstart[vox + 1]; ++k) {
                const int32_t cell = M.gcells[k];
                if (check_cell<D>(M, cell, testp)) {
                    found = cell;
                    break;
                }
            }
        }
        for (int32_t k = 0; found < 0 && k < M.nbig; ++k)
            if (check_cell<D>(M, M.big[k], testp)) found = M.big[k];
    }
    if (found < 0) {
        const int32_t k = atomicAdd(&sc->mesh_lost, 1);
        if (lost_list) lost_list[k] = int32_t(i);
        return;
    }
    s.nfail[i] = 0;
    if (!same) s.internal[i] = 0;
    s.cellID[i] = found;
#pragma unroll
    for (int a = 0; a < D; ++a) s.cellV[i * D + a] = M.cVel[(int64_t)found * D + a];
    s.cellP[i] = M.cP[found];
    s.cellRho[i] = M.cRho[found];
}

/* The mesh field at arbitrary points: the containing cell (-1 if none, with
 * zero values) found over the voxel candidates, vals = (vel[D], p, rho). */
template <int D>
__global__ __launch_bounds__(TPB) void k_mesh_eval(DevMesh M, int64_t n, const double* __restrict__ pts,
                                                   int32_t* __restrict__ cell_out, double* __restrict__ vals)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    double testp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) testp[a] = pts[i * D + a];
    int32_t found = -1;
    int64_t vox = 0;
    bool inside = true;
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
        const double q = (testp[a] - M.glo[a]) * M.ginv;
        if (!(q >= 0.0) || q >= double(M.gdims[a])) inside = false;
        const int32_t g = inside ? int32_t(q) : 0;
        vox = vox * M.gdims[a] + g;
    }
    if (inside) {
        for (int32_t k = M.gstart[vox]; k < M.gstart[vox + 1]; ++k) {
            const int32_t cell = M.gcells[k];
            if (check_cell<D>(M, cell, testp)) {
                found = cell;
                break;
            }
        }
    }
    for (int32_t k = 0; found < 0 && k < M.nbig; ++k)
        if (check_cell<D>(M, M.big[k], testp)) found = M.big[k];
    cell_out[i] = found;
#pragma unroll
    for (int a = 0; a < D; ++a) vals[i * (D + 2) + a] = found >= 0 ? M.cVel[(int64_t)found * D + a] : 0.0;
    vals[i * (D + 2) + D] = found >= 0 ? M.cP[found] : 0.0;
    vals[i * (D + 2) + D + 1] = found >= 0 ? M.cRho[found] : 0.0;
}

/* ------------------------------------------------ particles in no cell
 *
 * FindCell's boundary test (src/Containment.cpp:734-805) for a particle
 * that no cell contains, one block per particle:
 *
 * 1. the KNN_K = 500 nearest cell centres, nearest first (the reference's
 *    KD-tree query; equal distances here by cell index).  With fewer cells
 *    the reference's result vector keeps its zero-initialised tail, so
 *    cell 0 fills the list.  A radius holding between KNN_K and KNN_CAP
 *    centres is found by bisection over the centre grid; those centres are
 *    sorted in LDS by (squared distance, index);
 * 2. CheckCell over the list: the reference's 5- and 500-nearest searches
 *    stop at the first containing cell in this order;
 * 3. for each listed cell, the segment from the particle to the cell's
 *    centre against each of the cell's boundary faces (right cell < 0) in
 *    face order (Crossings3D / get_line_intersection); a hit flips the
 *    crossing parity, and a wall face (-1) sets internal or a far-field face
 *    (-2) marks the particle for deletion, either ending that cell's loop;
 * 4. with even parity ipt_n_failed counts up, or once it is past 10 the
 *    particle is deleted.
 */
constexpr int KNN_K = 500;
constexpr int KNN_CAP = 4096;

/* get_line_intersection (src/Geometry.cpp:312-341): does the segment
 * p1-cellC cross the edge?  Only edges with denominator >= machine epsilon
 * count (the reference's test is one-sided). */
__device__ int line_intersection(const DevMesh& M, int32_t f, const double* p1, const double* cellC)
{
#pragma clang fp contract(off)
    const int32_t* fv = M.face_v + M.face_off[f];
    const double* e1 = mvert(M, fv[0]);
    const double* e2 = mvert(M, fv[1]);
    const double s0 = cellC[0] - p1[0], s1 = cellC[1] - p1[1];
    const double r0 = e2[0] - e1[0], r1 = e2[1] - e1[1];
    const double denom = (-r0 * s1 + s0 * r1);
    if (denom < __DBL_EPSILON__) return 0;
    const double u = (-s1 * (p1[0] - e1[0]) + s0 * (p1[1] - e1[1])) / denom;
    const double t = (r0 * (p1[1] - e1[1]) - r1 * (p1[0] - e1[0])) / denom;
    return (u > 0 && u < 1 && t > 0 && t < 1) ? 1 : 0;
}

template <int D>
__device__ __forceinline__ double centre_d2(const DevMesh& M, int32_t cell, const double* p)
{
#pragma clang fp contract(off)
    double d2 = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double t = M.cCentre[(int64_t)cell * 3 + a] - p[a];
        d2 += t * t;
    }
    return d2;
}

/* Voxels of the centre grid within r of p along each axis, one voxel of
 * margin; false if none. */
template <int D>
__device__ bool centre_range(const DevMesh& M, const double* p, double r, int32_t* lo, int32_t* hi)
{
    lo[0] = lo[1] = lo[2] = 0;
    hi[0] = hi[1] = hi[2] = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const double ql = (p[a] - r - M.clo[a]) * M.cinv - 1.0, qh = (p[a] + r - M.clo[a]) * M.cinv + 1.0;
        if (!(qh >= 0.0) || !(ql < double(M.cdims[a]))) return false;
        lo[a] = ql <= 0.0 ? 0 : int32_t(floor(ql));
        hi[a] = qh >= double(M.cdims[a] - 1) ? M.cdims[a] - 1 : int32_t(floor(qh));
    }
    return true;
}

/* f(cell) for every centre in the voxel box, voxels dealt to the threads. */
template <typename F>
__device__ void for_centres(const DevMesh& M, const int32_t* lo, const int32_t* hi, F&& f)
{
    const int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const int64_t nv = nx * ny * nz;
    for (int64_t v = threadIdx.x; v < nv; v += TPB) {
        const int64_t ix = lo[0] + v % nx, iy = lo[1] + (v / nx) % ny, iz = lo[2] + v / (nx * ny);
        const int64_t vox = (iz * M.cdims[1] + iy) * M.cdims[0] + ix;
        for (int32_t k = M.cstart[vox]; k < M.cstart[vox + 1]; ++k) f(M.ccells[k]);
    }
}

/* The K nearest cell centres of p, nearest first (equal distances by cell
 * index), into list[0..K) (cell 0 past the mesh's cell count): a radius
 * holding between min(K, ncells) and KNN_CAP centres found by bisection
 * over the centre grid, those centres sorted in LDS.  Every thread of the
 * block calls it; false if no such radius was found. */
template <int D>
__device__ bool nearest_list(const DevMesh& M, const double* p, int32_t K, double* key, int32_t* idx, int32_t* list,
                             int32_t* s_cnt)
{
    const int32_t kt = M.ncells < K ? M.ncells : K;

    /* a radius with kt..KNN_CAP centres (all of them if they fit) */
    double r = INFINITY, rlo = 0.0, rhi = INFINITY;
    if (M.ncells > KNN_CAP) {
        double out2 = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const double lo = M.clo[a], hi = M.clo[a] + double(M.cdims[a]) / M.cinv;
            const double e = p[a] < lo ? lo - p[a] : (p[a] > hi ? p[a] - hi : 0.0);
            out2 += e * e;
        }
        r = sqrt(out2) + 0.5 * pow(double(kt), 1.0 / D) / M.cinv;
    }
    bool ok = false;
    for (int it = 0; it < 200 && !ok; ++it) {
        if (threadIdx.x == 0) *s_cnt = 0;
        __syncthreads();
        int32_t lo[3], hi[3], mine = 0;
        const double r2 = r * r;
        if (centre_range<D>(M, p, r, lo, hi))
            for_centres(M, lo, hi, [&](int32_t cell) { mine += centre_d2<D>(M, cell, p) <= r2 ? 1 : 0; });
        if (mine) atomicAdd(s_cnt, mine);
        __syncthreads();
        const int32_t cnt = *s_cnt;
        __syncthreads();
        if (cnt < kt) {
            rlo = r;
            r = isinf(rhi) ? 2.0 * r : 0.5 * (rlo + rhi);
        } else if (cnt > KNN_CAP) {
            rhi = r;
            r = 0.5 * (rlo + rhi);
        } else
            ok = true;
    }
    if (!ok) return false;

    /* the centres within r, sorted by (distance, index) */
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
    {
        int32_t lo[3], hi[3];
        const double r2 = r * r;
        if (centre_range<D>(M, p, r, lo, hi))
            for_centres(M, lo, hi, [&](int32_t cell) {
                const double d2 = centre_d2<D>(M, cell, p);
                if (d2 <= r2) {
                    const int32_t k = atomicAdd(s_cnt, 1);
                    if (k < KNN_CAP) {
                        key[k] = d2;
                        idx[k] = cell;
                    }
                }
            });
    }
    __syncthreads();
    const int32_t cnt = *s_cnt;
    __syncthreads();
    if (cnt < kt || cnt > KNN_CAP) return false; /* the count above was made with the same test */
    int32_t P = 1;
    while (P < cnt) P <<= 1;
    for (int32_t k = cnt + threadIdx.x; k < P; k += TPB) {
        key[k] = INFINITY;
        idx[k] = INT32_MAX;
    }
    __syncthreads();
    for (int32_t k = 2; k <= P; k <<= 1)
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            for (int32_t i = threadIdx.x; i < P; i += TPB) {
                const int32_t l = i ^ j;
                if (l > i) {
                    const bool gt = key[i] > key[l] || (key[i] == key[l] && idx[i] > idx[l]);
                    if (gt == ((i & k) == 0)) {
                        const double tk = key[i];
                        key[i] = key[l];
                        key[l] = tk;
                        const int32_t ti = idx[i];
                        idx[i] = idx[l];
                        idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    for (int32_t j = threadIdx.x; j < K; j += TPB) list[j] = j < kt ? idx[j] : 0;
    __syncthreads();
    return true;
}

template <int D>
__global__ __launch_bounds__(TPB) void k_lost_classify(DevState s, DevMesh M, const int32_t* __restrict__ lost_list,
                                                       int32_t* __restrict__ lost_del, DevScalars* sc)
{
    __shared__ double key[KNN_CAP];
    __shared__ int32_t idx[KNN_CAP];
    __shared__ int32_t list[KNN_K];
    __shared__ int32_t s_cnt, s_first, s_par, s_wall, s_far;
    const int32_t slot = lost_list[blockIdx.x];
    double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) p[a] = s.x[(int64_t)slot * D + a];

    /* 1. */
    if (!nearest_list<D>(M, p, KNN_K, key, idx, list, &s_cnt)) {
        if (threadIdx.x == 0) {
            lost_del[blockIdx.x] = 0;
            atomicOr(&sc->lost_err, 1);
        }
        return;
    }
    if (threadIdx.x == 0) {
        s_first = KNN_K;
        s_par = s_wall = s_far = 0;
    }
    __syncthreads();

    /* 2. the first listed cell that contains the particle */
    for (int32_t j = threadIdx.x; j < KNN_K; j += TPB)
        if (check_cell<D>(M, list[j], p)) atomicMin(&s_first, j);
    __syncthreads();
    if (s_first < KNN_K) {
        if (threadIdx.x == 0) {
            const int32_t cell = list[s_first];
            s.cellID[slot] = cell;
#pragma unroll
            for (int a = 0; a < D; ++a) s.cellV[(int64_t)slot * D + a] = M.cVel[(int64_t)cell * D + a];
            s.cellP[slot] = M.cP[cell];
            s.cellRho[slot] = M.cRho[cell];
            s.nfail[slot] = 0;
            s.internal[slot] = 0;
            lost_del[blockIdx.x] = 0;
            atomicSub(&sc->mesh_lost, 1);
        }
        return;
    }

    /* 3. boundary crossings on the way to each listed centre */
    int par = 0, wall = 0, far = 0;
    for (int32_t j = threadIdx.x; j < KNN_K; j += TPB) {
        const int32_t cell = list[j];
        const double* rayp = M.cCentre + (int64_t)cell * 3;
        for (int32_t k = M.cell_off[cell]; k < M.cell_off[cell + 1]; ++k) {
            const int32_t f = M.cell_f[k];
            const int32_t right = M.face_lr[2 * (int64_t)f + 1];
            if (right >= 0) continue;
            const int hit = D == 3 ? crossings3d(M, f, p, rayp) : line_intersection(M, f, p, rayp);
            if (!hit) continue;
            par ^= 1;
            if (right == -1) {
                wall = 1;
                break;
            }
            if (right == -2) {
                far = 1;
                break;
            }
        }
    }
    if (par) atomicXor(&s_par, 1);
    if (wall) atomicOr(&s_wall, 1);
    if (far) atomicOr(&s_far, 1);
    __syncthreads();

    /* 4. */
    if (threadIdx.x == 0) {
        bool del = s_far != 0;
        if (s_wall) {
            s.internal[slot] = 1;
            atomicAdd(&sc->lost_int, 1);
        }
        if (!s_par) {
            atomicAdd(&sc->lost_fail, 1);
            if (s.nfail[slot] > 10) del = true;
            else s.nfail[slot] += 1;
        }
        lost_del[blockIdx.x] = del ? 1 : 0;
        if (del) atomicAdd(&sc->lost_del, 1);
    }
}

/* FirstCell for a particle leaving the pipe (one block each): if the
 * lookup before gave it no cell that holds it, the KNN_FIRST nearest cells
 * (the reference's 150 in 3D, 20 in 2D) and CheckCell over them; if none
 * holds it, the boundary branch in the reference's order, in one thread:
 * for every listed cell and each of its boundary faces in face order, the
 * segment from the particle's position before this test to the cell's
 * centre; every hit flips the parity, a wall face reflects the velocity
 * off the face's normal (cross product of its first two edges, normalised;
 * in 2D the edge turned left) and moves the particle onto the face's plane
 * (through its second vertex), a far-field face marks it for deletion.  No
 * crossing at all is the reference's fatal error (lost_err 2). */
template <int D>
__global__ __launch_bounds__(TPB) void k_first_cell(DevState s, DevMesh M, const int32_t* __restrict__ fresh,
                                                    int32_t* __restrict__ fc_del, DevScalars* sc)
{
#pragma clang fp contract(off)
    constexpr int32_t K = D == 3 ? 150 : 20;
    __shared__ double key[KNN_CAP];
    __shared__ int32_t idx[KNN_CAP];
    __shared__ int32_t list[K];
    __shared__ int32_t s_cnt, s_first;
    const int32_t slot = fresh[blockIdx.x];
    if (threadIdx.x == 0) fc_del[blockIdx.x] = 0;
    double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < D; ++a) p[a] = s.x[(int64_t)slot * D + a];
    const int32_t cur = s.cellID[slot];
    if (cur >= 0 && cur < M.ncells && check_cell<D>(M, cur, p)) return; /* found by the lookup */
    if (!nearest_list<D>(M, p, K, key, idx, list, &s_cnt)) {
        if (threadIdx.x == 0) atomicOr(&sc->lost_err, 1);
        return;
    }
    if (threadIdx.x == 0) s_first = K;
    __syncthreads();
    for (int32_t j = threadIdx.x; j < K; j += TPB)
        if (check_cell<D>(M, list[j], p)) atomicMin(&s_first, j);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (s_first < K) {
        const int32_t cell = list[s_first];
        s.cellID[slot] = cell;
#pragma unroll
        for (int a = 0; a < D; ++a) s.cellV[(int64_t)slot * D + a] = M.cVel[(int64_t)cell * D + a];
        s.cellP[slot] = M.cP[cell];
        s.cellRho[slot] = M.cRho[cell];
        return;
    }
    double xi[D], v[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        xi[a] = p[a];
        v[a] = s.v[(int64_t)slot * D + a];
    }
    int cross = 0;
    bool del = false;
    for (int32_t j = 0; j < K; ++j) {
        const int32_t cell = list[j];
        const double* rayp = M.cCentre + (int64_t)cell * 3;
        for (int32_t k = M.cell_off[cell]; k < M.cell_off[cell + 1]; ++k) {
            const int32_t f = M.cell_f[k];
            const int32_t right = M.face_lr[2 * (int64_t)f + 1];
            if (right >= 0) continue;
            const int hit = D == 3 ? crossings3d(M, f, p, rayp) : line_intersection(M, f, p, rayp);
            if (!hit) continue;
            cross ^= 1;
            if (right == -1) {
                const int32_t* fv = M.face_v + M.face_off[f];
                double nrm[D];
                if (D == 3) {
                    const double* v0 = mvert(M, fv[0]);
                    const double* v1 = mvert(M, fv[1]);
                    const double* v2 = mvert(M, fv[2]);
                    const double r1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
                    const double r2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
                    nrm[0] = r1[1] * r2[2] - r1[2] * r2[1];
                    nrm[1] = r1[2] * r2[0] - r1[0] * r2[2];
                    nrm[D - 1] = r1[0] * r2[1] - r1[1] * r2[0];
                } else {
                    const double* v0 = mvert(M, fv[0]);
                    const double* v1 = mvert(M, fv[1]);
                    nrm[0] = -(v1[1] - v0[1]);
                    nrm[1] = v1[0] - v0[0];
                }
                double z = 0.0;
#pragma unroll
                for (int a = 0; a < D; ++a) z += nrm[a] * nrm[a];
                if (z > 0.0) {
                    const double rn = sqrt(z);
#pragma unroll
                    for (int a = 0; a < D; ++a) nrm[a] = nrm[a] / rn;
                }
                double vn = 0.0, xn = 0.0, nn = 0.0, plane = 0.0;
                const double* q = mvert(M, fv[1]);
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    vn += v[a] * nrm[a];
                    plane += nrm[a] * q[a];
                }
#pragma unroll
                for (int a = 0; a < D; ++a) v[a] = v[a] - (2 * vn) * nrm[a];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    xn += xi[a] * nrm[a];
                    nn += nrm[a] * nrm[a];
                }
                const double dist = (plane - xn) / nn;
#pragma unroll
                for (int a = 0; a < D; ++a) xi[a] = xi[a] + dist * nrm[a];
            } else if (right == -2) {
                del = true;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        s.x[(int64_t)slot * D + a] = xi[a];
        s.v[(int64_t)slot * D + a] = v[a];
    }
    atomicAdd(&sc->fc_count, 1);
    if (!cross) atomicOr(&sc->lost_err, 2);
    if (del) {
        fc_del[blockIdx.x] = 1;
        atomicAdd(&sc->fc_del, 1);
    }
}

void launch_first_cell(int dim, DevState s, DevMesh m, const int32_t* fresh, int32_t nfresh, int32_t* fc_del,
                       DevScalars* sc, hipStream_t st)
{
    if (nfresh <= 0) return;
    const dim3 g(static_cast<unsigned>(nfresh)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_first_cell<2>, g, b, 0, st, s, m, fresh, fc_del, sc);
    else hipLaunchKernelGGL(k_first_cell<3>, g, b, 0, st, s, m, fresh, fc_del, sc);
}

__global__ __launch_bounds__(TPB) void k_keep_all(int32_t* __restrict__ keep, int32_t* __restrict__ hflag, int64_t n)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    keep[i] = 1;
    hflag[i] = 0;
}

__global__ __launch_bounds__(TPB) void k_lost_marks(const int32_t* __restrict__ lost_list,
                                                    const int32_t* __restrict__ lost_del, int32_t nlost,
                                                    const int32_t* __restrict__ part_id,
                                                    const int32_t* __restrict__ blk, int32_t* __restrict__ keep,
                                                    int32_t* __restrict__ hflag, int32_t* __restrict__ blkcount)
{
    const int32_t k = blockIdx.x * TPB + threadIdx.x;
    if (k >= nlost || !lost_del[k]) return;
    const int32_t slot = lost_list[k];
    keep[slot] = 0;
    hflag[part_id[slot]] = 1;
    if (blk[slot] >= 0) atomicAdd(blkcount + blk[slot], 1);
}

void launch_lost_classify(int dim, DevState s, DevMesh m, const int32_t* lost_list, int32_t nlost, int32_t* lost_del,
                          DevScalars* sc, hipStream_t st)
{
    if (nlost <= 0) return;
    const dim3 g(static_cast<unsigned>(nlost)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_lost_classify<2>, g, b, 0, st, s, m, lost_list, lost_del, sc);
    else hipLaunchKernelGGL(k_lost_classify<3>, g, b, 0, st, s, m, lost_list, lost_del, sc);
}

void launch_lost_marks(const int32_t* lost_list, const int32_t* lost_del, int32_t nlost, const int32_t* part_id,
                       const int32_t* blk, int64_t n, int32_t* keep, int32_t* hflag, int32_t* blkcount,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_keep_all, dim3(grid_of(n)), dim3(TPB), 0, st, keep, hflag, n);
    if (nlost > 0)
        hipLaunchKernelGGL(k_lost_marks, dim3(grid_of(nlost)), dim3(TPB), 0, st, lost_list, lost_del, nlost, part_id,
                           blk, keep, hflag, blkcount);
}

void launch_mesh_eval(int dim, DevMesh m, int64_t n, const double* pts, int32_t* cell, double* vals, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_mesh_eval<2>, g, b, 0, st, m, n, pts, cell, vals);
    else hipLaunchKernelGGL(k_mesh_eval<3>, g, b, 0, st, m, n, pts, cell, vals);
}

void launch_find_cell(int dim, const Consts& c, DevState s, const int32_t* part_id, DevMesh m, int64_t n,
                      DevScalars* sc, int32_t* lost_list, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_find_cell<2>, g, b, 0, st, c, s, part_id, m, n, sc, lost_list);
    else hipLaunchKernelGGL(k_find_cell<3>, g, b, 0, st, c, s, part_id, m, n, sc, lost_list);
}

}  // namespace fj
