/*
 * TAU mesh and solution ingestion (src/CDFIO.cpp) into the flat mesh arrays
 * of fjsph_set_mesh.  Host code.
 */
#include "fjsph_tau.h"

#include <cmath>
#include <fstream>
#include <string>
#include <vector>

#include "netcdf.hpp"

struct fjsph_tau_mesh {
    int dim = 3, ignored = 0;
    std::vector<double> verts;
    std::vector<int64_t> face_off{0}, face_verts, face_lr, cell_off, cell_faces;
    std::vector<double> vel, p, rho;
    std::vector<int64_t> surf;
    std::vector<int32_t> markers;
    int64_t n_sol = 0;
};

namespace {

thread_local std::string g_err;

struct TauError {
    fjsph_status code;
    std::string msg;
};

[[noreturn]] void fail(const std::string& m, fjsph_status c = FJSPH_ERR_STATE)
{
    throw TauError{c, m};
}

/* Get_Coordinates (src/CDFIO.cpp:462-628): in 2D exactly one of
 * points_xc/yc/zc is missing and names the dropped axis. */
void read_coordinates(fj::nc::File& f, int dim, uint64_t npts, fjsph_tau_mesh& m)
{
    const char* names[3] = {"points_xc", "points_yc", "points_zc"};
    std::vector<int> present;
    for (int a = 0; a < 3; ++a)
        if (f.has_var(names[a])) present.push_back(a);
    std::vector<std::vector<double>> c;
    if (dim == 3) {
        if (present.size() != 3) fail("mesh file lacks a coordinate variable");
        m.ignored = 0;
    } else {
        if (present.size() != 2) fail("2D mesh file must hold exactly two of points_xc/yc/zc");
        m.ignored = 1 + (present[0] == 0 ? (present[1] == 1 ? 2 : 1) : 0);
    }
    for (int a : present) {
        c.push_back(f.get<double>(names[a]));
        if (c.back().size() != npts) fail(std::string(names[a]) + " does not hold no_of_points values");
    }
    m.verts.resize(size_t(npts) * size_t(dim));
    for (uint64_t i = 0; i < npts; ++i)
        for (int k = 0; k < dim; ++k) m.verts[size_t(i) * dim + k] = c[size_t(k)][size_t(i)];
}

std::vector<int32_t> ints(fj::nc::File& f, const std::string& name, uint64_t n)
{
    std::vector<int32_t> v = f.get<int32_t>(name);
    if (v.size() != n) fail(name + ": expected " + std::to_string(n) + " values, found " + std::to_string(v.size()));
    return v;
}

/* Place_Faces / Place_Edges (src/CDFIO.cpp:830-990, 1103-1226): left/right
 * per face, each cell's faces in face order, boundary faces (right < 0). */
void place_faces(fj::nc::File& f, const char* left_name, const char* right_name, uint64_t nface, uint64_t nelem,
                 fjsph_tau_mesh& m)
{
    const std::vector<int32_t> left = ints(f, left_name, nface), right = ints(f, right_name, nface);
    m.face_lr.resize(size_t(nface) * 2);
    std::vector<int64_t> count(size_t(nelem) + 1, 0);
    for (uint64_t i = 0; i < nface; ++i) {
        const int32_t l = left[size_t(i)], r = right[size_t(i)];
        if (l < 0 || uint64_t(l) >= nelem || (r >= 0 && uint64_t(r) >= nelem))
            fail("face " + std::to_string(i) + " names an element out of range");
        m.face_lr[size_t(i) * 2] = l;
        m.face_lr[size_t(i) * 2 + 1] = r;
        count[size_t(l) + 1]++;
        if (r >= 0) count[size_t(r) + 1]++;
        else m.surf.push_back(int64_t(i));
    }
    for (size_t c = 0; c < size_t(nelem); ++c) count[c + 1] += count[c];
    m.cell_off = count;
    m.cell_faces.assign(size_t(count.back()), 0);
    std::vector<int64_t> at(count.begin(), count.end() - 1);
    for (uint64_t i = 0; i < nface; ++i) {
        m.cell_faces[size_t(at[size_t(left[size_t(i)])]++)] = int64_t(i);
        if (right[size_t(i)] >= 0) m.cell_faces[size_t(at[size_t(right[size_t(i)])]++)] = int64_t(i);
    }
    if (f.has_var("boundarymarker_of_surfaces")) {
        const std::vector<int32_t> mk = f.get<int32_t>("boundarymarker_of_surfaces");
        if (mk.size() != m.surf.size()) fail("boundarymarker_of_surfaces does not match the boundary faces");
        m.markers = mk;
    } else {
        for (int64_t s : m.surf) m.markers.push_back(int32_t(m.face_lr[size_t(s) * 2 + 1]));
    }
}

void append_faces(fj::nc::File& f, const std::string& var, uint64_t nf, uint64_t per, uint64_t npts, fjsph_tau_mesh& m)
{
    const std::vector<int32_t> v = ints(f, var, nf * per);
    for (uint64_t i = 0; i < nf; ++i) {
        for (uint64_t k = 0; k < per; ++k) {
            const int32_t q = v[size_t(i * per + k)];
            if (q < 0 || uint64_t(q) >= npts) fail(var + ": vertex out of range");
            m.face_verts.push_back(q);
        }
        m.face_off.push_back(int64_t(m.face_verts.size()));
    }
}

/* Read_tau_mesh_EDGE (src/CDFIO.cpp:992-1100).  no_of_surfaceelements and
 * boundarymarker_of_surfaces are optional here (files of the older
 * converter hold no_of_wall_edges / no_of_farfield_edges instead). */
std::vector<int32_t> read_edge_mesh(const std::string& path, fjsph_tau_mesh& m)
{
    fj::nc::File f(path);
    const uint64_t nelem = f.dim_len("no_of_elements");
    const uint64_t nedge = f.dim_len("no_of_edges");
    const uint64_t ppe = f.dim_len("points_per_edge");
    const uint64_t npts = f.dim_len("no_of_points");
    if (ppe != 2) fail("points_per_edge must be 2");
    append_faces(f, "points_of_element_edges", nedge, ppe, npts, m);
    std::vector<int32_t> used;
    if (f.has_var("vertices_in_use")) used = ints(f, "vertices_in_use", npts);
    read_coordinates(f, 2, npts, m);
    place_faces(f, "left_element_of_edges", "right_element_of_edges", nedge, nelem, m);
    if (f.has_dim("no_of_surfaceelements") && f.dim_len("no_of_surfaceelements") != m.surf.size())
        fail("no_of_surfaceelements does not match the boundary edges");
    return used;
}

/* Read_tau_mesh_FACE (src/CDFIO.cpp:1228-1356) */
void read_face_mesh(const std::string& path, fjsph_tau_mesh& m)
{
    fj::nc::File f(path);
    const uint64_t nelem = f.dim_len("no_of_elements");
    const uint64_t nface = f.dim_len("no_of_faces");
    const uint64_t npts = f.dim_len("no_of_points");
    if (f.has_dim("no_of_triangles"))
        append_faces(f, "points_of_triangles", f.dim_len("no_of_triangles"), f.dim_len("points_per_triangle"), npts, m);
    if (f.has_dim("no_of_quadrilaterals"))
        append_faces(f, "points_of_quadrilaterals", f.dim_len("no_of_quadrilaterals"),
                     f.dim_len("points_per_quadrilateral"), npts, m);
    if (m.face_off.size() - 1 != nface) fail("the triangles and quadrilaterals do not add up to no_of_faces");
    read_coordinates(f, 3, npts, m);
    place_faces(f, "left_element_of_faces", "right_element_of_faces", nface, nelem, m);
    if (f.has_dim("no_of_surfaceelements") && f.dim_len("no_of_surfaceelements") != m.surf.size())
        fail("no_of_surfaceelements does not match the boundary faces");
}

/* Read_SOLUTION (src/CDFIO.cpp:655-822) */
void read_solution(const std::string& path, int offset_axis, const std::vector<int32_t>& used, fjsph_tau_mesh& m)
{
    fj::nc::File f(path);
    const uint64_t n = f.dim_len("no_of_points");
    m.n_sol = int64_t(n);
    const size_t nv = m.verts.size() / size_t(m.dim);
    auto field = [&](const char* name) {
        std::vector<double> v = f.get<double>(name);
        if (v.size() != n) fail(std::string(name) + " does not hold no_of_points values");
        return v;
    };
    std::vector<std::vector<double>> comp;
    if (m.dim == 3) {
        if (n != nv) fail("solution and mesh have different point counts");
        comp = {field("x_velocity"), field("y_velocity"), field("z_velocity")};
    } else {
        if (offset_axis == 1) comp = {field("y_velocity"), field("z_velocity")};
        else if (offset_axis == 2) comp = {field("x_velocity"), field("z_velocity")};
        else if (offset_axis == 3) comp = {field("x_velocity"), field("y_velocity")};
        else fail("2D offset axis must be 1, 2 or 3", FJSPH_ERR_ARG);
    }
    const std::vector<double> press = field("pressure"), dens = field("density");
    /* the solution point of each mesh vertex: through vertices_in_use when
     * the mesh has it (2D), else the same index */
    std::vector<int64_t> src(nv);
    if (!used.empty()) {
        for (size_t i = 0; i < nv; ++i) {
            if (used[i] < 0 || uint64_t(used[i]) >= n) fail("vertices_in_use names a point out of the solution");
            src[i] = used[i];
        }
    } else {
        if (n != nv) fail("solution and mesh have different point counts");
        for (size_t i = 0; i < nv; ++i) src[i] = int64_t(i);
    }
    m.vel.resize(nv * size_t(m.dim));
    m.p.resize(nv);
    m.rho.resize(nv);
    for (size_t i = 0; i < nv; ++i) {
        for (int k = 0; k < m.dim; ++k) m.vel[i * m.dim + k] = comp[size_t(k)][size_t(src[i])];
        m.p[i] = press[size_t(src[i])];
        m.rho[i] = dens[size_t(src[i])];
    }
}

template <class F>
fjsph_status guarded(F&& f)
{
    try {
        f();
        return FJSPH_OK;
    } catch (const TauError& e) {
        g_err = e.msg;
        return e.code;
    } catch (const fj::nc::Error& e) {
        g_err = e.msg;
        return FJSPH_ERR_STATE;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return FJSPH_ERR_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return FJSPH_ERR_STATE;
    }
}

std::string ltrim(const std::string& s)
{
    size_t i = 0;
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
    return s.substr(i);
}

}  // namespace

extern "C" {

const char* fjsph_tau_last_error(void)
{
    return g_err.c_str();
}

fjsph_status fjsph_tau_open(const char* mesh_path, const char* sol_path, int32_t dim, int32_t offset_axis,
                            double scale, fjsph_tau_mesh** out)
{
    if (!mesh_path || !out || (dim != 2 && dim != 3)) return FJSPH_ERR_ARG;
    *out = nullptr;
    return guarded([&] {
        fjsph_tau_mesh* m = new fjsph_tau_mesh();
        try {
            m->dim = dim;
            std::vector<int32_t> used;
            if (dim == 2) used = read_edge_mesh(mesh_path, *m);
            else read_face_mesh(mesh_path, *m);
            for (double& v : m->verts) v *= scale; /* cells.scale (src/CDFIO.cpp:1080-1085) */
            const size_t nv = m->verts.size() / size_t(dim);
            if (sol_path) {
                read_solution(sol_path, offset_axis, used, *m);
            } else {
                m->vel.assign(nv * size_t(dim), 0.0);
                m->p.assign(nv, 0.0);
                m->rho.assign(nv, 0.0);
            }
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

void fjsph_tau_close(fjsph_tau_mesh* m)
{
    delete m;
}

fjsph_status fjsph_tau_get_sizes(const fjsph_tau_mesh* m, fjsph_tau_sizes* out)
{
    if (!m || !out) return FJSPH_ERR_ARG;
    out->dim = m->dim;
    out->ignored_axis = m->ignored;
    out->nverts = int64_t(m->verts.size() / size_t(m->dim));
    out->nfaces = int64_t(m->face_off.size() - 1);
    out->nface_verts = int64_t(m->face_verts.size());
    out->ncells = int64_t(m->cell_off.size() - 1);
    out->ncell_faces = int64_t(m->cell_faces.size());
    out->nsurf = int64_t(m->surf.size());
    out->n_sol_points = m->n_sol;
    return FJSPH_OK;
}

fjsph_status fjsph_tau_get_arrays(const fjsph_tau_mesh* m, double* verts, int64_t* face_off, int64_t* face_verts,
                                  int64_t* face_lr, int64_t* cell_off, int64_t* cell_faces, double* vert_vel,
                                  double* vert_p, double* vert_rho)
{
    if (!m) return FJSPH_ERR_ARG;
    auto put = [](const auto& v, auto* dst) {
        if (dst) std::copy(v.begin(), v.end(), dst);
    };
    put(m->verts, verts);
    put(m->face_off, face_off);
    put(m->face_verts, face_verts);
    put(m->face_lr, face_lr);
    put(m->cell_off, cell_off);
    put(m->cell_faces, cell_faces);
    put(m->vel, vert_vel);
    put(m->p, vert_p);
    put(m->rho, vert_rho);
    return FJSPH_OK;
}

fjsph_status fjsph_tau_get_surfaces(const fjsph_tau_mesh* m, int64_t* faces, int32_t* markers)
{
    if (!m) return FJSPH_ERR_ARG;
    if (faces) std::copy(m->surf.begin(), m->surf.end(), faces);
    if (markers) std::copy(m->markers.begin(), m->markers.end(), markers);
    return FJSPH_OK;
}

fjsph_status fjsph_set_mesh_tau(fjsph_ctx* ctx, const fjsph_tau_mesh* m)
{
    if (!ctx || !m) return FJSPH_ERR_ARG;
    return fjsph_set_mesh(ctx, int64_t(m->verts.size() / size_t(m->dim)), m->verts.data(),
                          int64_t(m->face_off.size() - 1), m->face_off.data(), m->face_verts.data(), m->face_lr.data(),
                          int64_t(m->cell_off.size() - 1), m->cell_off.data(), m->cell_faces.data(), m->vel.data(),
                          m->p.data(), m->rho.data());
}

fjsph_status fjsph_tau_read_bmap(const char* bmap_path, fjsph_settings* s, double* angle_alpha)
{
    if (!bmap_path) return FJSPH_ERR_ARG;
    return guarded([&] {
        std::ifstream in(bmap_path);
        if (!in.is_open()) fail(std::string("cannot open ") + bmap_path);
        double alpha = 0.0;
        std::string line;
        while (std::getline(in, line)) {
            line = ltrim(line);
            if (!line.empty() && line[0] == '#') continue;
            const size_t pos = line.find(':');
            if (pos == std::string::npos) continue;
            if (ltrim(line.substr(0, pos)) != "Angle alpha (degree)") continue;
            alpha = std::strtod(line.c_str() + pos + 1, nullptr);
        }
        alpha *= 3.14159265358979323846 / 180.0;
        if (angle_alpha) *angle_alpha = alpha;
        if (s) {
            /* Read_BMAP (src/CDFIO.cpp:302-308): the old y component feeds x */
            const double g1 = s->grav[1], g2 = s->grav[2];
            if (s->dim == 3) s->grav[2] = g2 * std::cos(alpha);
            else s->grav[1] = g1 * std::cos(alpha);
            s->grav[0] = -g1 * std::sin(alpha);
        }
    });
}

}  // extern "C"
