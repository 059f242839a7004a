/*
 * C entry points of the FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * They mirror the engine's C-ABI (include/fjsph_engine.h) with an orc_
 * prefix so tests can drive both with the same arrays, pass by pass.
 */
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "fjsph_oracle.hpp"

using namespace orc;

namespace {

struct Handle {
    int dim;
    std::unique_ptr<Sim<2>> s2;
    std::unique_ptr<Sim<3>> s3;
};

template <int D>
Sim<D>& sim(Handle* h);
template <>
Sim<2>& sim<2>(Handle* h) { return *h->s2; }
template <>
Sim<3>& sim<3>(Handle* h) { return *h->s3; }

template <int D>
int field_width(int field)
{
    switch (field) {
    case FJSPH_F_X: case FJSPH_F_V: case FJSPH_F_ACC: case FJSPH_F_AF: case FJSPH_F_AVISC:
    case FJSPH_F_GRADRHO: case FJSPH_F_NORM: case FJSPH_F_CELLV: case FJSPH_F_VPERT:
        return D;
    case FJSPH_F_L:
        return D * D;
    default:
        return 1;
    }
}

template <int D>
double* vec_ptr(Part<D>& q, int field)
{
    switch (field) {
    case FJSPH_F_X: return q.xi.c;
    case FJSPH_F_V: return q.v.c;
    case FJSPH_F_ACC: return q.acc.c;
    case FJSPH_F_AF: return q.Af.c;
    case FJSPH_F_AVISC: return q.aVisc.c;
    case FJSPH_F_GRADRHO: return q.gradRho.c;
    case FJSPH_F_NORM: return q.norm.c;
    case FJSPH_F_CELLV: return q.cellV.c;
    case FJSPH_F_VPERT: return q.vPert.c;
    case FJSPH_F_L: return &q.L.a[0][0];
    case FJSPH_F_RHO: return &q.rho;
    case FJSPH_F_P: return &q.p;
    case FJSPH_F_RRHO: return &q.Rrho;
    case FJSPH_F_M: return &q.m;
    case FJSPH_F_DELTAD: return &q.deltaD;
    case FJSPH_F_LAM: return &q.lam;
    case FJSPH_F_LAM_NB: return &q.lam_nb;
    case FJSPH_F_COLOURG: return &q.colourG;
    case FJSPH_F_COLOUR: return &q.colour;
    case FJSPH_F_KERNSUM: return &q.kernsum;
    case FJSPH_F_CURVE: return &q.curve;
    case FJSPH_F_NORM_CURVE: return &q.norm_curve;
    case FJSPH_F_WOCCL: return &q.woccl;
    case FJSPH_F_PDIST: return &q.pDist;
    case FJSPH_F_CELLP: return &q.cellP;
    case FJSPH_F_CELLRHO: return &q.cellRho;
    default: return nullptr;
    }
}

template <int D>
int download(Handle* h, int which, int field, double* out)
{
    Sim<D>& S = sim<D>(h);
    std::vector<Part<D>>& st = which == FJSPH_STATE_N ? S.pn : S.pnp1;
    const int w = field_width<D>(field);
    for (size_t i = 0; i < st.size(); ++i) {
        Part<D>& q = st[i];
        switch (field) {
        case FJSPH_F_SURF: out[i] = q.surf; break;
        case FJSPH_F_CELLID: out[i] = double(q.cellID); break;
        case FJSPH_F_SURFZONE: out[i] = q.surfzone; break;
        case FJSPH_F_B: out[i] = q.b; break;
        case FJSPH_F_PID: out[i] = double(q.part_id); break;
        case FJSPH_F_NFAILED: out[i] = q.n_failed; break;
        case FJSPH_F_INTERNAL: out[i] = q.internal; break;
        default: {
            double* p = vec_ptr<D>(q, field);
            if (!p) return FJSPH_ERR_ARG;
            for (int k = 0; k < w; ++k) out[i * w + k] = p[k];
        }
        }
    }
    return FJSPH_OK;
}

template <int D>
int upload(Handle* h, int which, int field, const double* in)
{
    Sim<D>& S = sim<D>(h);
    std::vector<Part<D>>& st = which == FJSPH_STATE_N ? S.pn : S.pnp1;
    const int w = field_width<D>(field);
    for (size_t i = 0; i < st.size(); ++i) {
        Part<D>& q = st[i];
        switch (field) {
        case FJSPH_F_SURF: q.surf = uint32_t(in[i]); break;
        case FJSPH_F_CELLID: q.cellID = long(in[i]); break;
        case FJSPH_F_SURFZONE: q.surfzone = uint32_t(in[i]); break;
        case FJSPH_F_B: q.b = uint32_t(in[i]); break;
        case FJSPH_F_PID: q.part_id = int64_t(in[i]); break;
        case FJSPH_F_NFAILED: q.n_failed = uint32_t(in[i]); break;
        case FJSPH_F_INTERNAL: q.internal = uint32_t(in[i]); break;
        default: {
            double* p = vec_ptr<D>(q, field);
            if (!p) return FJSPH_ERR_ARG;
            for (int k = 0; k < w; ++k) p[k] = in[i * w + k];
        }
        }
    }
    return FJSPH_OK;
}

template <int D>
void upload_state(Handle* h, int64_t n, int64_t nb, const double* x, const double* v, const double* rho,
                  const double* p, const double* m, const int32_t* b)
{
    Sim<D>& S = sim<D>(h);
    S.pn.assign(size_t(n), Part<D>());
    for (int64_t i = 0; i < n; ++i) {
        Part<D>& q = S.pn[i];
        for (int k = 0; k < D; ++k) { q.xi.c[k] = x[i * D + k]; q.v.c[k] = v[i * D + k]; }
        q.rho = rho[i];
        q.p = p[i];
        q.m = m[i];
        q.b = uint32_t(b[i]);
        q.part_id = i;
        q.cellID = kNoCell;
        if (S.s.aero_source != FJSPH_MESH_INFL) {
            q.cellRho = S.s.rho_g;
            q.cellP = S.s.p_ref;
            for (int k = 0; k < D; ++k) q.cellV.c[k] = S.s.v_inf[k];
        }
    }
    S.pnp1 = S.pn;
    S.bound_points = size_t(nb);
    S.total_points = size_t(n);
    S.start_index = S.bound_points;
    S.end_index = S.total_points;
    S.current_time = 0.0;
    S.last_frame_time = 0.0;
    S.cfl = S.s.cfl;
    S.delta_t = S.s.delta_t_min > 0 ? S.s.delta_t_min : 2e-10;
    S.n_stable = S.n_unstable = 0;
    S.current_frame = 0;
    S.part_id_next = size_t(n); /* src/Init.cpp:478 */
    S.outlist.clear();
}

void fill_info(fjsph_step_info* info, double rms, double dt, double cfl, double t, double mrp, double mf, double maf,
               double npd, unsigned it, int rb)
{
    if (!info) return;
    info->rms_error = rms;
    info->dt = dt;
    info->cfl = cfl;
    info->time = t;
    info->max_rho_pc = mrp;
    info->maxf = mf;
    info->maxAf = maf;
    info->npd = npd;
    info->iterations = int32_t(it);
    info->rollbacks = rb;
}

template <int D>
int step(Handle* h, int kind, fjsph_step_info* info)
{
    Sim<D>& S = sim<D>(h);
    double rms = 0.0;
    try {
        rms = kind == 0 ? S.integrate_no_update() : S.integrate();
    } catch (const std::exception&) {
        /* a fatal condition of the reference (it prints and exits) */
        return FJSPH_ERR_STATE;
    }
    fill_info(info, rms, S.delta_t, S.cfl, S.current_time, S.maxRho_pc, S.maxf, S.maxAf, S.last_npd, S.iteration, S.rollbacks);
    return FJSPH_OK;
}

template <int D>
int set_mesh_t(Handle* h, int64_t nverts, const double* verts, int64_t nfaces, const int64_t* face_off,
               const int64_t* face_verts, const int64_t* face_lr, int64_t ncells, const int64_t* cell_off,
               const int64_t* cell_faces, const double* vvel, const double* vp, const double* vrho)
{
    Sim<D>& S = sim<D>(h);
    MeshData& M = S.mesh;
    M.dim = D;
    M.verts.assign(verts, verts + nverts * D);
    M.face_off.assign(face_off, face_off + nfaces + 1);
    M.face_v.assign(face_verts, face_verts + face_off[nfaces]);
    M.face_l.resize(size_t(nfaces));
    M.face_r.resize(size_t(nfaces));
    for (int64_t f = 0; f < nfaces; ++f) { M.face_l[size_t(f)] = face_lr[2 * f]; M.face_r[size_t(f)] = face_lr[2 * f + 1]; }
    M.cell_off.assign(cell_off, cell_off + ncells + 1);
    M.cell_f.assign(cell_faces, cell_faces + cell_off[ncells]);
    average_point_data_to_cell(M, vvel, vp, vrho);
    S.have_mesh = true;
    return FJSPH_OK;
}

void to_public(const orc::IptRecord& a, fjsph_ipt_record& b)
{
    b.part_id = a.pid;
    b.status = a.status;
    b.steps = a.steps;
    b.cell = a.cell;
    b.face = a.face;
    b.t0 = a.t0;
    b.t = a.t;
    for (int k = 0; k < 3; ++k) {
        b.x0[k] = a.x0[k];
        b.x[k] = a.x[k];
        b.v[k] = a.v[k];
    }
}

template <int D>
int ipt_track_t(Handle* h, int64_t n, const double* x, const double* v, const int32_t* cell, const int64_t* pid,
                double t0, fjsph_ipt_record* out)
{
    Sim<D>& S = sim<D>(h);
    if (!S.have_mesh) return FJSPH_ERR_STATE;
    const orc::IptParams P = S.ipt_params();
    const orc::MeshData& M = S.mesh;
    std::vector<orc::IptRecord> rec(static_cast<size_t>(n));
#pragma omp parallel for num_threads(S.nthreads) schedule(dynamic, 4)
    for (int64_t i = 0; i < n; ++i) {
        orc::IptP<D> q;
        q.pid = pid ? pid[i] : i;
        q.t = t0;
        q.cellID = cell[i];
        for (int a = 0; a < D; ++a) {
            q.xi.c[a] = x[i * D + a];
            q.v.c[a] = v[i * D + a];
        }
        if (q.cellID >= 0 && q.cellID < int64_t(M.ncells())) {
            for (int a = 0; a < D; ++a) q.cellV.c[a] = M.cVel[size_t(q.cellID) * D + a];
            q.cellRho = M.cRho[size_t(q.cellID)];
        }
        q.d = P.diam;
        q.A = P.area;
        rec[size_t(i)] = orc::ipt_integrate<D>(M, P, q);
    }
    for (int64_t i = 0; i < n; ++i) to_public(rec[size_t(i)], out[i]);
    return FJSPH_OK;
}

}  // namespace

#define DISPATCH(h, call) ((h)->dim == 2 ? call<2> : call<3>)

extern "C" {

void* orc_create(const fjsph_settings* s, int nthreads)
{
    if (!s || (s->dim != 2 && s->dim != 3)) return nullptr;
    Handle* h = new Handle();
    h->dim = s->dim;
    if (s->dim == 2) {
        h->s2.reset(new Sim<2>());
        h->s2->s = *s;
        h->s2->d = derive(*s);
        h->s2->nthreads = nthreads > 0 ? nthreads : 1;
    } else {
        h->s3.reset(new Sim<3>());
        h->s3->s = *s;
        h->s3->d = derive(*s);
        h->s3->nthreads = nthreads > 0 ? nthreads : 1;
    }
    return h;
}

void orc_destroy(void* hp) { delete static_cast<Handle*>(hp); }

int orc_derive(const fjsph_settings* s, fjsph_derived* out)
{
    if (!s || !out) return FJSPH_ERR_ARG;
    *out = derive(*s);
    return FJSPH_OK;
}

int orc_set_serial_dissipation(void* hp, int on)
{
    Handle* h = static_cast<Handle*>(hp);
    if (h->dim == 2) h->s2->serial_dissipation = on; else h->s3->serial_dissipation = on;
    return FJSPH_OK;
}

int orc_upload_state(void* hp, int64_t n, int64_t nb, const double* x, const double* v, const double* rho,
                     const double* p, const double* m, const int32_t* b)
{
    Handle* h = static_cast<Handle*>(hp);
    if (n <= 0 || nb < 0 || nb > n) return FJSPH_ERR_ARG;
    DISPATCH(h, upload_state)(h, n, nb, x, v, rho, p, m, b);
    return FJSPH_OK;
}

int orc_set_blocks(void* hp, int nblocks, const fjsph_block* blocks)
{
    Handle* h = static_cast<Handle*>(hp);
    std::vector<Block> bl(static_cast<size_t>(nblocks));
    for (int i = 0; i < nblocks; ++i) bl[i].b = blocks[i];
    if (h->dim == 2) h->s2->blocks = bl; else h->s3->blocks = bl;
    return FJSPH_OK;
}

int orc_set_block_motion(void* hp, int block, int n_times, const double* times, const double* vels)
{
    Handle* h = static_cast<Handle*>(hp);
    std::vector<Block>& bl = h->dim == 2 ? h->s2->blocks : h->s3->blocks;
    if (block < 0 || block >= int(bl.size())) return FJSPH_ERR_ARG;
    bl[block].b.n_times = n_times;
    bl[block].times.assign(times, times + n_times);
    bl[block].vels.assign(vels, vels + 3 * n_times);
    return FJSPH_OK;
}

int orc_set_block_planes(void* hp, int block, const fjsph_block_planes* pl)
{
    Handle* h = static_cast<Handle*>(hp);
    std::vector<Block>& bl = h->dim == 2 ? h->s2->blocks : h->s3->blocks;
    if (!pl || block < 0 || block >= int(bl.size())) return FJSPH_ERR_ARG;
    const int inlet = bl[block].pl.inlet;
    bl[block].pl = *pl;
    bl[block].pl.inlet = pl->inlet || inlet;
    return FJSPH_OK;
}

/* back[c], buffer[c*n_buf + j]: host indices (bound_block::back/buffer) */
int orc_set_inlet(void* hp, int block, int n_cols, int n_buf, const int64_t* back, const int64_t* buffer)
{
    Handle* h = static_cast<Handle*>(hp);
    std::vector<Block>& bl = h->dim == 2 ? h->s2->blocks : h->s3->blocks;
    if (block < 0 || block >= int(bl.size()) || n_cols <= 0 || n_buf <= 0) return FJSPH_ERR_ARG;
    Block& b = bl[block];
    b.back.assign(back, back + n_cols);
    b.buffer.assign(size_t(n_cols), std::vector<size_t>());
    for (int c = 0; c < n_cols; ++c) b.buffer[c].assign(buffer + size_t(c) * n_buf, buffer + size_t(c + 1) * n_buf);
    b.pl.inlet = 1;
    return FJSPH_OK;
}

int orc_get_inlet(void* hp, int block, int* n_cols, int* n_buf, int64_t* back, int64_t* buffer)
{
    Handle* h = static_cast<Handle*>(hp);
    std::vector<Block>& bl = h->dim == 2 ? h->s2->blocks : h->s3->blocks;
    if (block < 0 || block >= int(bl.size())) return FJSPH_ERR_ARG;
    const Block& b = bl[block];
    *n_cols = int(b.back.size());
    *n_buf = b.buffer.empty() ? 0 : int(b.buffer[0].size());
    if (!back || !buffer) return FJSPH_OK;
    for (size_t c = 0; c < b.back.size(); ++c) {
        back[c] = int64_t(b.back[c]);
        for (size_t j = 0; j < b.buffer[c].size(); ++j) buffer[c * size_t(*n_buf) + j] = int64_t(b.buffer[c][j]);
    }
    return FJSPH_OK;
}

int orc_set_max_points(void* hp, int64_t max_points)
{
    Handle* h = static_cast<Handle*>(hp);
    if (h->dim == 2) h->s2->max_points = size_t(max_points); else h->s3->max_points = size_t(max_points);
    return FJSPH_OK;
}

int orc_get_counts(void* hp, int64_t* n_total, int64_t* n_bound)
{
    Handle* h = static_cast<Handle*>(hp);
    *n_total = int64_t(h->dim == 2 ? h->s2->pnp1.size() : h->s3->pnp1.size());
    *n_bound = int64_t(h->dim == 2 ? h->s2->bound_points : h->s3->bound_points);
    return FJSPH_OK;
}

int orc_get_block_range(void* hp, int block, int64_t* first, int64_t* last)
{
    Handle* h = static_cast<Handle*>(hp);
    const std::vector<Block>& bl = h->dim == 2 ? h->s2->blocks : h->s3->blocks;
    if (block < 0 || block >= int(bl.size())) return FJSPH_ERR_ARG;
    *first = bl[block].b.first;
    *last = bl[block].b.last;
    return FJSPH_OK;
}

int orc_download_field(void* hp, int which, int field, double* out)
{
    Handle* h = static_cast<Handle*>(hp);
    return DISPATCH(h, download)(h, which, field, out);
}

int orc_upload_field(void* hp, int which, int field, const double* in)
{
    Handle* h = static_cast<Handle*>(hp);
    return DISPATCH(h, upload)(h, which, field, in);
}

int orc_build_neighbours(void* hp, int brute)
{
    Handle* h = static_cast<Handle*>(hp);
    if (h->dim == 2) {
        auto& S = *h->s2;
        S.outlist = brute ? brute_neighbours<2>(S.pnp1, S.d.sr) : find_neighbours<2>(S.pnp1, S.d.sr, S.nthreads);
    } else {
        auto& S = *h->s3;
        S.outlist = brute ? brute_neighbours<3>(S.pnp1, S.d.sr) : find_neighbours<3>(S.pnp1, S.d.sr, S.nthreads);
    }
    return FJSPH_OK;
}

int64_t orc_neighbour_total(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    const OUTL& o = h->dim == 2 ? h->s2->outlist : h->s3->outlist;
    int64_t t = 0;
    for (const auto& l : o) t += int64_t(l.size());
    return t;
}

/* CSR dump: offsets[n+1], idx/dist2[total]; order as stored (sorted). */
int orc_get_neighbours(void* hp, int64_t* offsets, int64_t* idx, double* dist2)
{
    Handle* h = static_cast<Handle*>(hp);
    const OUTL& o = h->dim == 2 ? h->s2->outlist : h->s3->outlist;
    int64_t t = 0;
    for (size_t i = 0; i < o.size(); ++i) {
        offsets[i] = t;
        for (const auto& e : o[i]) {
            idx[t] = e.first;
            if (dist2) dist2[t] = e.second;
            ++t;
        }
    }
    offsets[o.size()] = t;
    return FJSPH_OK;
}

int orc_dsph_prestep(void* hp, double* npd)
{
    Handle* h = static_cast<Handle*>(hp);
    double v = 1.0;
    if (h->dim == 2) h->s2->dSPH_PreStep(h->s2->total_points, v); else h->s3->dSPH_PreStep(h->s3->total_points, v);
    if (npd) *npd = v;
    return FJSPH_OK;
}

#define SIMCALL(h, expr2, expr3) do { if ((h)->dim == 2) { auto& S = *(h)->s2; expr2; } else { auto& S = *(h)->s3; expr3; } } while (0)

int orc_aero_velocity(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    double npd = 1.0;
    SIMCALL(h, S.get_aero_velocity(S.bound_points, S.total_points, npd),
            S.get_aero_velocity(S.bound_points, S.total_points, npd));
    return FJSPH_OK;
}

int orc_set_mesh(void* hp, int64_t nverts, const double* verts, int64_t nfaces, const int64_t* face_off,
                 const int64_t* face_verts, const int64_t* face_lr, int64_t ncells, const int64_t* cell_off,
                 const int64_t* cell_faces, const double* vvel, const double* vp, const double* vrho)
{
    Handle* h = static_cast<Handle*>(hp);
    return h->dim == 2 ? set_mesh_t<2>(h, nverts, verts, nfaces, face_off, face_verts, face_lr, ncells, cell_off,
                                       cell_faces, vvel, vp, vrho)
                       : set_mesh_t<3>(h, nverts, verts, nfaces, face_off, face_verts, face_lr, ncells, cell_off,
                                       cell_faces, vvel, vp, vrho);
}

int orc_get_mesh_cells(void* hp, double* centre, double* vel, double* p, double* rho)
{
    Handle* h = static_cast<Handle*>(hp);
    const MeshData& M = h->dim == 2 ? h->s2->mesh : h->s3->mesh;
    std::copy(M.cCentre.begin(), M.cCentre.end(), centre);
    std::copy(M.cVel.begin(), M.cVel.end(), vel);
    std::copy(M.cP.begin(), M.cP.end(), p);
    std::copy(M.cRho.begin(), M.cRho.end(), rho);
    return FJSPH_OK;
}

int orc_mesh_lost(void* hp, int64_t* lost)
{
    Handle* h = static_cast<Handle*>(hp);
    *lost = h->dim == 2 ? h->s2->mesh_lost : h->s3->mesh_lost;
    return FJSPH_OK;
}

/* lost, deleted, internal, failed, first_cell, first_cell_deleted
 * (fjsph_lost_info). */
int orc_mesh_lost_info(void* hp, int64_t* out6)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, (out6[0] = S.mesh_lost, out6[1] = S.mesh_deleted, out6[2] = S.mesh_internal, out6[3] = S.mesh_failed,
                out6[4] = S.first_cell_reflected, out6[5] = S.first_cell_deleted),
            (out6[0] = S.mesh_lost, out6[1] = S.mesh_deleted, out6[2] = S.mesh_internal, out6[3] = S.mesh_failed,
             out6[4] = S.first_cell_reflected, out6[5] = S.first_cell_deleted));
    return FJSPH_OK;
}

/* FindCell's boundary test of one point against its 500-nearest list:
 * out3 = (cross, wall, far) (for pinning tests). */
int orc_boundary_test(void* hp, const double* point, int32_t* out3)
{
    Handle* h = static_cast<Handle*>(hp);
    const MeshData& M = h->dim == 2 ? h->s2->mesh : h->s3->mesh;
    const BoundaryTest r = boundary_test(M, point, knn_cells_padded(M, point, 500));
    out3[0] = r.cross;
    out3[1] = r.wall ? 1 : 0;
    out3[2] = r.far ? 1 : 0;
    return FJSPH_OK;
}

/* FirstCell's boundary branch for one point and velocity (pinning tests):
 * xv = (x[3], v[3]) in and out; out2 = (cross, del). */
}  // extern "C"
template <int D>
int first_cell_t(Handle* h, double* xv, int32_t* out2)
{
    const MeshData& M = sim<D>(h).mesh;
    Vec<D> x, v;
    for (int a = 0; a < D; ++a) {
        x.c[a] = xv[a];
        v.c[a] = xv[3 + a];
    }
    const std::vector<int64_t> cand = knn_cells_padded(M, x.c, D == 3 ? 150 : 20);
    const FirstCellResult r = first_cell_boundary<D>(M, cand, x, v);
    for (int a = 0; a < D; ++a) {
        xv[a] = x.c[a];
        xv[3 + a] = v.c[a];
    }
    out2[0] = r.cross;
    out2[1] = r.del ? 1 : 0;
    return FJSPH_OK;
}
extern "C" {

int orc_first_cell_boundary(void* hp, double* xv, int32_t* out2)
{
    Handle* h = static_cast<Handle*>(hp);
    return h->dim == 2 ? first_cell_t<2>(h, xv, out2) : first_cell_t<3>(h, xv, out2);
}

/* get_line_intersection of one segment against one 2D edge. */
int orc_line_intersection(void* hp, int64_t face, const double* p1, const double* p2)
{
    Handle* h = static_cast<Handle*>(hp);
    const MeshData& M = h->dim == 2 ? h->s2->mesh : h->s3->mesh;
    return line_intersection(M, face, p1, p2);
}

/* CheckCell of one point against one cell (for pinning tests). */
int orc_check_cell(void* hp, int64_t cell, const double* point)
{
    Handle* h = static_cast<Handle*>(hp);
    const MeshData& M = h->dim == 2 ? h->s2->mesh : h->s3->mesh;
    return check_cell(M, cell, point);
}

int orc_detect_surface(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, S.Detect_Surface(S.bound_points, S.total_points), S.Detect_Surface(S.bound_points, S.total_points));
    return FJSPH_OK;
}

int orc_dissipation_terms(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, S.dissipation_terms(S.bound_points, S.total_points), S.dissipation_terms(S.bound_points, S.total_points));
    return FJSPH_OK;
}

int orc_particle_shift(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, S.particle_shift(S.bound_points, S.total_points), S.particle_shift(S.bound_points, S.total_points));
    return FJSPH_OK;
}

int orc_boundary_update(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, { std::vector<int> ni(S.bound_points, 0); S.boundary_update(S.pnp1, ni, false); },
               { std::vector<int> ni(S.bound_points, 0); S.boundary_update(S.pnp1, ni, false); });
    return FJSPH_OK;
}

int orc_acc_rrho(void* hp, double npd)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, S.get_acc_and_Rrho(npd, S.pnp1), S.get_acc_and_Rrho(npd, S.pnp1));
    return FJSPH_OK;
}

int orc_nb_iter(void* hp, double npd, double* log_error)
{
    Handle* h = static_cast<Handle*>(hp);
    double le = 0.0;
    SIMCALL(h, { S.start_index = S.bound_points; S.end_index = S.total_points; S.save_xih(); S.Do_NB_Iter(npd); le = S.log_error_of(S.xih); },
               { S.start_index = S.bound_points; S.end_index = S.total_points; S.save_xih(); S.Do_NB_Iter(npd); le = S.log_error_of(S.xih); });
    if (log_error) *log_error = le;
    return FJSPH_OK;
}

int orc_find_timestep(void* hp, double* dt)
{
    Handle* h = static_cast<Handle*>(hp);
    double v = 0.0;
    SIMCALL(h, { S.start_index = S.bound_points; S.end_index = S.total_points; v = S.find_timestep(); },
               { S.start_index = S.bound_points; S.end_index = S.total_points; v = S.find_timestep(); });
    if (dt) *dt = v;
    return FJSPH_OK;
}

int orc_integrate_no_update(void* hp, fjsph_step_info* info)
{
    Handle* h = static_cast<Handle*>(hp);
    return DISPATCH(h, step)(h, 0, info);
}

int orc_integrate_step(void* hp, fjsph_step_info* info)
{
    Handle* h = static_cast<Handle*>(hp);
    return DISPATCH(h, step)(h, 1, info);
}

int orc_advance_frame(void* hp)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, { S.last_frame_time += S.s.frame_time_interval; S.current_frame++; },
            { S.last_frame_time += S.s.frame_time_interval; S.current_frame++; });
    return FJSPH_OK;
}

extern "C++" {
template <class SimT>
void get_clock(const SimT& S, fjsph_clock* c)
{
    c->current_time = S.current_time;
    c->last_frame_time = S.last_frame_time;
    c->delta_t = S.delta_t;
    c->cfl = S.cfl;
    c->current_frame = S.current_frame;
    c->n_stable = S.n_stable;
    c->n_unstable = S.n_unstable;
    c->next_part_id = int64_t(S.part_id_next);
}

template <class SimT>
void set_clock(SimT& S, const fjsph_clock* c)
{
    S.current_time = c->current_time;
    S.last_frame_time = c->last_frame_time;
    S.delta_t = c->delta_t;
    S.cfl = c->cfl;
    S.current_frame = unsigned(c->current_frame);
    S.n_stable = unsigned(c->n_stable);
    S.n_unstable = unsigned(c->n_unstable);
    S.part_id_next = size_t(c->next_part_id);
}
}  // extern "C++"

int orc_set_ipt(void* hp, const fjsph_ipt_settings* s)
{
    Handle* h = static_cast<Handle*>(hp);
    if (!s) return FJSPH_ERR_ARG;
    if (h->dim == 2) h->s2->ipt = *s; else h->s3->ipt = *s;
    return FJSPH_OK;
}

int orc_ipt_count(void* hp, int64_t* n, int64_t* ok, int64_t* failed)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, { *n = int64_t(S.ipt_records.size()); *ok = S.ipt_success; *failed = S.ipt_failed; },
            { *n = int64_t(S.ipt_records.size()); *ok = S.ipt_success; *failed = S.ipt_failed; });
    return FJSPH_OK;
}

int orc_ipt_records(void* hp, fjsph_ipt_record* out)
{
    Handle* h = static_cast<Handle*>(hp);
    const std::vector<orc::IptRecord>& r = h->dim == 2 ? h->s2->ipt_records : h->s3->ipt_records;
    for (size_t i = 0; i < r.size(); ++i) to_public(r[i], out[i]);
    return FJSPH_OK;
}

int orc_ipt_track(void* hp, int64_t n, const double* x, const double* v, const int32_t* cell, const int64_t* pid,
                  double t0, fjsph_ipt_record* out)
{
    Handle* h = static_cast<Handle*>(hp);
    return h->dim == 2 ? ipt_track_t<2>(h, n, x, v, cell, pid, t0, out) : ipt_track_t<3>(h, n, x, v, cell, pid, t0, out);
}

/* The reference's ray-triangle test (MollerTrumbore, src/Geometry.cpp:668-716)
 * on one triangle, for unit checks: returns 1 when it reports a crossing. */
int orc_moller(const double* v0, const double* v1, const double* v2, const double* origin, const double* dir,
               double* dt, double* denom)
{
    return orc::moller(orc::v3(v0), orc::v3(v1), orc::v3(v2), orc::v3(origin), orc::v3(dir), *dt, *denom) ? 1 : 0;
}

int orc_get_clock(void* hp, fjsph_clock* c)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, get_clock(S, c), get_clock(S, c));
    return FJSPH_OK;
}

int orc_set_clock(void* hp, const fjsph_clock* c)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, set_clock(S, c), set_clock(S, c));
    return FJSPH_OK;
}

int orc_get_time(void* hp, double* t, double* dt, double* cfl)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, { *t = S.current_time; *dt = S.delta_t; *cfl = S.cfl; }, { *t = S.current_time; *dt = S.delta_t; *cfl = S.cfl; });
    return FJSPH_OK;
}

int orc_set_time(void* hp, double t, double dt, double cfl)
{
    Handle* h = static_cast<Handle*>(hp);
    SIMCALL(h, { S.current_time = t; S.delta_t = dt; S.cfl = cfl; }, { S.current_time = t; S.delta_t = dt; S.cfl = cfl; });
    return FJSPH_OK;
}

/* Unit access to the Eigen restatements (row-major input). */
int orc_qr_inverse(int dim, const double* m, double* inv)
{
    if (dim == 2) {
        Mat<2> a, r;
        for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) a.a[i][j] = m[i * 2 + j];
        if (!colpiv_qr_inverse<2>(a, r)) return 0;
        for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) inv[i * 2 + j] = r.a[i][j];
        return 1;
    }
    Mat<3> a, r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a.a[i][j] = m[i * 3 + j];
    if (!colpiv_qr_inverse<3>(a, r)) return 0;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) inv[i * 3 + j] = r.a[i][j];
    return 1;
}

double orc_min_eig(int dim, const double* m)
{
    if (dim == 2) {
        Mat<2> a;
        for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) a.a[i][j] = m[i * 2 + j];
        return min_eig_direct(a);
    }
    Mat<3> a;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a.a[i][j] = m[i * 3 + j];
    return min_eig_direct(a);
}

/* Pointwise kernels for closed-form checks. */
double orc_kernel(double r, double H, double Wc) { return Kernel(r, H, Wc); }
double orc_gradk_scalar(double r, double H, double Wc)
{
    Vec<1> one;
    one.c[0] = 1.0;
    return GradK<1>(one, r, H, Wc).c[0];
}

}  // extern "C"
