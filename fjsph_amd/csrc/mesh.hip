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
                                                   DevMesh M, int64_t n, int32_t* lost)
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
    if (prev >= 0 && prev < M.ncells && check_cell<D>(M, prev, testp)) found = prev;
    if (found < 0) {
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
        atomicAdd(lost, 1);
        return;
    }
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

void launch_mesh_eval(int dim, DevMesh m, int64_t n, const double* pts, int32_t* cell, double* vals, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_mesh_eval<2>, g, b, 0, st, m, n, pts, cell, vals);
    else hipLaunchKernelGGL(k_mesh_eval<3>, g, b, 0, st, m, n, pts, cell, vals);
}

void launch_find_cell(int dim, const Consts& c, DevState s, const int32_t* part_id, DevMesh m, int64_t n,
                      int32_t* lost, hipStream_t st)
{
    const dim3 g(grid_of(n)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_find_cell<2>, g, b, 0, st, c, s, part_id, m, n, lost);
    else hipLaunchKernelGGL(k_find_cell<3>, g, b, 0, st, c, s, part_id, m, n, lost);
}

}  // namespace fj
