/*
 * Restart files in the reference's HDF5 layout (Write_HDF5 / Read_HDF5,
 * src/H5IO.cpp:915-1318).  Host code: a state image (struct Image) is moved
 * between the file, a caller's host arrays (fjsph_restart_*_host) and an
 * engine context (fjsph_restart_write / fjsph_restart_read, through the
 * engine's C-ABI).
 *
 * HDF5 types follow the reference's writers (src/H5IO.cpp:49-130, 189-380):
 * size_t values as NATIVE_ULONG (u64), uint and int values as NATIVE_UINT
 * (u32, the int writer also uses the unsigned type), reals as IEEE_F64LE,
 * vectors as a 1-D f64 attribute of dim values, strings as space-padded
 * Fortran strings of their length (empty strings are not written).
 * Datasets: reals f64, "Boundary condition" u32, "Particle ID" u64,
 * "Cell ID" i64 (long), "Back index" / "Buffer index" u64.
 *
 * One deviation from the reference reader: its Read_Attribute reads every
 * scalar through a size_t temporary and converts that integer to the
 * target type (src/H5IO.cpp:571-585), so real attributes (current time,
 * CFL, spacing) come back as the integer value of their bit pattern.  This
 * reader reads reals as reals.
 */
#include "fjsph_restart.h"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct RestartError {
    fjsph_status code;
    std::string msg;
};

[[noreturn]] void fail(const std::string& m, fjsph_status code = FJSPH_ERR_STATE)
{
    throw RestartError{code, m};
}

/* One HDF5 handle, closed by the function that fits its kind. */
class H5 {
public:
    H5(hid_t id, herr_t (*close)(hid_t), const std::string& what) : id_(id), close_(close)
    {
        if (id_ < 0) fail("HDF5: " + what + " failed");
    }
    H5(const H5&) = delete;
    H5& operator=(const H5&) = delete;
    ~H5()
    {
        if (id_ >= 0) close_(id_);
    }
    operator hid_t() const { return id_; }

private:
    hid_t id_;
    herr_t (*close_)(hid_t);
};

void check(herr_t e, const std::string& what)
{
    if (e < 0) fail("HDF5: " + what + " failed");
}

/* ------------------------------------------------------------ state image */

struct Image {
    int dim = 3;
    int64_t n = 0, nb = 0;
    fjsph_settings s{};
    fjsph_clock clock{};
    int64_t max_points = 9999999;
    std::vector<fjsph_block> blocks;
    std::vector<fjsph_block_planes> planes;
    std::vector<std::vector<double>> times, vels; /* vels: 3 per time */
    std::vector<std::vector<int64_t>> back, buffer;
    std::vector<double> x, v, acc, p, rho, rrho, m, cell_v, cell_p, cell_rho;
    std::vector<int32_t> b;
    std::vector<int64_t> pid, cell_id;
    /* root attributes a reader takes from the file */
    double particle_step = 0.0, frame_time_interval = 0.0, scale = 1.0;
    int32_t single_file = 1;
};

const char* dim_name(int d)
{
    return d == 0 ? "x" : (d == 1 ? "y" : "z");
}

/* ---------------------------------------------------------------- writing */

void attr_raw(hid_t loc, const std::string& name, hid_t ftype, hid_t mtype, const void* v)
{
    H5 sp(H5Screate(H5S_SCALAR), H5Sclose, "H5Screate");
    H5 a(H5Acreate2(loc, name.c_str(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "attribute " + name);
    check(H5Awrite(a, mtype, v), "writing attribute " + name);
}

void attr_u64(hid_t loc, const std::string& name, uint64_t v)
{
    unsigned long w = (unsigned long)v;
    attr_raw(loc, name, H5T_NATIVE_ULONG, H5T_NATIVE_ULONG, &w);
}

void attr_u32(hid_t loc, const std::string& name, uint32_t v)
{
    unsigned w = v;
    attr_raw(loc, name, H5T_NATIVE_UINT, H5T_NATIVE_UINT, &w);
}

/* Write_Int_Attribute stores an int under the unsigned type */
void attr_int(hid_t loc, const std::string& name, int32_t v)
{
    unsigned w;
    std::memcpy(&w, &v, sizeof(w));
    attr_raw(loc, name, H5T_NATIVE_UINT, H5T_NATIVE_UINT, &w);
}

void attr_f64(hid_t loc, const std::string& name, double v)
{
    attr_raw(loc, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &v);
}

void attr_vec(hid_t loc, const std::string& name, const double* v, int dim)
{
    hsize_t dims[1] = {hsize_t(dim)};
    H5 sp(H5Screate_simple(1, dims, nullptr), H5Sclose, "H5Screate_simple");
    H5 a(H5Acreate2(loc, name.c_str(), H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "attribute " + name);
    check(H5Awrite(a, H5T_NATIVE_DOUBLE, v), "writing attribute " + name);
}

void attr_str(hid_t loc, const std::string& name, const char* v)
{
    if (!v || !*v) return;
    const size_t len = std::strlen(v);
    H5 ft(H5Tcopy(H5T_FORTRAN_S1), H5Tclose, "H5Tcopy");
    check(H5Tset_size(ft, len), "H5Tset_size");
    H5 mt(H5Tcopy(H5T_C_S1), H5Tclose, "H5Tcopy");
    check(H5Tset_size(mt, len + 1), "H5Tset_size");
    H5 sp(H5Screate(H5S_SCALAR), H5Sclose, "H5Screate");
    H5 a(H5Acreate2(loc, name.c_str(), ft, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "attribute " + name);
    check(H5Awrite(a, mt, v), "writing attribute " + name);
}

void dset_raw(hid_t loc, const std::string& name, hid_t ftype, hid_t mtype, const void* data, size_t n)
{
    hsize_t dims[1] = {hsize_t(n)};
    H5 sp(H5Screate_simple(1, dims, nullptr), H5Sclose, "H5Screate_simple");
    H5 d(H5Dcreate2(loc, name.c_str(), ftype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
         "dataset " + name);
    if (n > 0) check(H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "writing " + name);
}

void dset_f64(hid_t loc, const std::string& name, const std::vector<double>& v)
{
    dset_raw(loc, name, H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, v.data(), v.size());
}

const char* solver_name_of(const fjsph_settings& s)
{
    return s.integrator == FJSPH_RUNGE_KUTTA ? "Runge-Kutta" : "Newmark-Beta";
}

const char* aero_name_of(const fjsph_settings& s)
{
    switch (s.aero_case) {
    case FJSPH_GISSLER: return "Gissler";
    case FJSPH_INDUCED_PRESSURE: return "Induced_pressure";
    case FJSPH_SKIN_FRIC: return "Skin_friction";
    default: return "(none)";
    }
}

/* Write_HDF5_Attributes (src/H5IO.cpp:915-1025), in its order. */
void write_root(hid_t f, const Image& im, const fjsph_restart_meta& mt, int n_bound_blocks, int n_fluid_blocks,
                int64_t bound_points, int64_t fluid_points)
{
    const fjsph_settings& s = im.s;
    const int D = im.dim;
    attr_f64(f, "Simulation current time", im.clock.current_time);
    attr_u64(f, "Particle index to add", uint64_t(im.clock.next_part_id));
    attr_u64(f, "Number of boundary blocks", uint64_t(n_bound_blocks));
    attr_u64(f, "Number of fluid blocks", uint64_t(n_fluid_blocks));
    attr_u64(f, "Number of boundary points", uint64_t(bound_points));
    attr_u64(f, "Number of fluid points", uint64_t(fluid_points));
    attr_u32(f, "Current frame", uint32_t(im.clock.current_frame));

    attr_str(f, "Input para filename", mt.input_file);
    attr_str(f, "Input fluid definition filename", mt.fluid_file);
    attr_str(f, "Input boundary definition filename", mt.bound_file);
    attr_str(f, "Primary grid face filename", mt.tau_mesh);
    attr_str(f, "Boundary mapping filename", mt.tau_bmap);
    attr_str(f, "Restart-data prefix", mt.tau_sol);
    attr_int(f, "Dimension offset vector", mt.offset_axis);
    attr_f64(f, "Solution angle of attack", mt.angle_alpha);
    attr_f64(f, "Grid scale", mt.scale);

    attr_str(f, "OpenFOAM input directory", mt.foam_dir);
    attr_str(f, "OpenFOAM solution directory", mt.foam_sol);
    attr_int(f, "OpenFOAM binary", mt.foam_is_binary);
    attr_int(f, "Label size", mt.foam_label_size);
    attr_int(f, "Scalar size", mt.foam_scalar_size);
    attr_int(f, "OpenFOAM is buoyant", mt.foam_buoyant_sim);
    attr_int(f, "OpenFOAM is foam_is_incompressible", mt.foam_is_incomp);

    attr_str(f, "VLM definition filename", mt.vlm_file);

    attr_u32(f, "Single file for output", uint32_t(mt.single_file));
    attr_str(f, "Output files prefix", mt.output_prefix);
    attr_str(f, "SPH restart prefix", mt.restart_prefix);
    attr_f64(f, "SPH frame time interval", s.frame_time_interval);
    attr_u32(f, "SPH frame count", uint32_t(mt.max_frames));
    attr_f64(f, "SPH previous frame time", im.clock.last_frame_time);
    attr_u32(f, "SPH output encoding", uint32_t(mt.out_encoding));
    attr_str(f, "Variable list", mt.output_names);

    attr_f64(f, "Reference density", s.rho_g);
    attr_f64(f, "Reference dispersed density", s.rho_rest);
    attr_f64(f, "Sutherland reference viscosity", s.mu_g);
    attr_f64(f, "Reference dispersed viscosity", s.mu);
    attr_f64(f, "Reference surface tension", s.sig);
    attr_f64(f, "SPH surface tension contact angle", mt.contangb);
    attr_int(f, "Init hydrostatic pressure", mt.init_hydro_pressure);
    attr_f64(f, "Hydrostatic height", mt.hydro_height);

    attr_f64(f, "Reference velocity", mt.v_ref);
    attr_f64(f, "Reference pressure", s.p_ref);
    attr_f64(f, "Reference Mach number", mt.M_ref);
    attr_f64(f, "Reference temperature", s.temp_g);
    attr_f64(f, "Gas constant gamma", s.gamma_g);

    attr_str(f, "SPH integration solver", mt.solver_name ? mt.solver_name : solver_name_of(s));
    attr_u32(f, "SPH boundary solver", uint32_t(mt.bound_solver));
    attr_f64(f, "SPH solver minimum residual", s.min_residual);
    attr_f64(f, "SPH maximum timestep", s.delta_t_max);
    attr_f64(f, "SPH minimum timestep", s.delta_t_min);
    attr_f64(f, "SPH maximum CFL", s.cfl_max);
    attr_f64(f, "SPH minimum CFL", s.cfl_min);
    attr_f64(f, "SPH CFL condition", im.clock.cfl);
    attr_f64(f, "SPH unstable CFL step", s.cfl_step);
    attr_u32(f, "SPH unstable CFL count limit", uint32_t(s.n_unstable_limit));
    attr_u32(f, "SPH stable CFL count limit", uint32_t(s.n_stable_limit));
    attr_f64(f, "SPH stable CFL count iteration factor", s.subits_factor);
    attr_f64(f, "SPH maximum shifting velocity", s.max_shift_vel);
    attr_u32(f, "SPH stable CFL count", uint32_t(im.clock.n_stable));
    attr_u32(f, "SPH unstable CFL count", uint32_t(im.clock.n_unstable));

    attr_f64(f, "SPH background pressure", s.press_back);
    attr_f64(f, "SPH starting pressure", s.press_pipe);
    attr_f64(f, "SPH density variation", s.rho_var);
    attr_f64(f, "SPH maximum density", s.rho_max);
    attr_f64(f, "SPH minimum density", s.rho_min);
    attr_f64(f, "SPH delta coefficient", s.dsph_delta);

    attr_f64(f, "SPH artificial viscosity factor", s.visc_alpha);
    attr_f64(f, "SPH speed of sound", s.speed_sound);
    attr_u32(f, "SPH Newmark Beta iteration limit", uint32_t(s.max_subits));
    attr_vec(f, "SPH gravity vector", s.grav, D);

    attr_f64(f, "SPH initial spacing", s.particle_step);
    attr_f64(f, "SPH boundary spacing factor", mt.bound_step_factor);
    attr_f64(f, "SPH smoothing length factor", s.H_fac);
    attr_str(f, "SPH aerodynamic case", mt.aero_case_name ? mt.aero_case_name : aero_name_of(s));
    attr_int(f, "SPH SP diameter definition", mt.use_dx);
    attr_int(f, "SPH use TAB deformation", s.use_TAB_def);
    attr_vec(f, "SPH global offset coordinate", mt.offset_vec, D);
    attr_u64(f, "SPH maximum particle count", uint64_t(im.max_points));
    attr_f64(f, "SPH aerodynamic cutoff value", s.lam_cutoff);
    attr_vec(f, "SPH freestream velocity", s.v_inf, D);
    attr_f64(f, "SPH restart fit tolerance", mt.restart_tol);

    attr_int(f, "Transition to IPT", mt.using_ipt);
    attr_int(f, "Velocity equation order", mt.ipt_eq_order);
    attr_f64(f, "SPH tracking conversion x coordinate", mt.max_x_sph);
    attr_f64(f, "Maximum x trajectory coordinate", mt.max_x);
    attr_u32(f, "Particle scatter output", uint32_t(mt.part_out));
    attr_u32(f, "Particle streak output", uint32_t(mt.streak_out));
    attr_u32(f, "Particle cell intersection output", uint32_t(mt.cells_out));
}

/* Write_Zone_Data (src/H5IO.cpp:395-523) for host range [first, last). */
void write_zone_data(hid_t g, const Image& im, double scale, int64_t first, int64_t last)
{
    const int D = im.dim;
    const size_t len = size_t(last - first);
    std::vector<double> vec(len);
    auto comp = [&](const std::vector<double>& a, int d, double div) {
        for (size_t i = 0; i < len; ++i) vec[i] = a[size_t(first + int64_t(i)) * D + d] / div;
    };
    auto scal = [&](const std::vector<double>& a) {
        for (size_t i = 0; i < len; ++i) vec[i] = a[size_t(first) + i];
    };
    for (int d = 0; d < D; ++d) { comp(im.x, d, scale); dset_f64(g, std::string("Position coordinate ") + dim_name(d), vec); }
    for (int d = 0; d < D; ++d) { comp(im.v, d, 1.0); dset_f64(g, std::string("Velocity ") + dim_name(d), vec); }
    for (int d = 0; d < D; ++d) { comp(im.acc, d, 1.0); dset_f64(g, std::string("Acceleration ") + dim_name(d), vec); }
    scal(im.p); dset_f64(g, "Pressure", vec);
    scal(im.rho); dset_f64(g, "Density", vec);
    scal(im.rrho); dset_f64(g, "Density gradient", vec);
    scal(im.m); dset_f64(g, "Mass", vec);
    std::vector<unsigned> uvec(len);
    for (size_t i = 0; i < len; ++i) uvec[i] = unsigned(im.b[size_t(first) + i]);
    dset_raw(g, "Boundary condition", H5T_NATIVE_UINT, H5T_NATIVE_UINT, uvec.data(), len);
    std::vector<unsigned long> svec(len);
    for (size_t i = 0; i < len; ++i) svec[i] = (unsigned long)im.pid[size_t(first) + i];
    dset_raw(g, "Particle ID", H5T_NATIVE_ULONG, H5T_NATIVE_ULONG, svec.data(), len);
    std::vector<long> lvec(len);
    for (size_t i = 0; i < len; ++i) lvec[i] = long(im.cell_id[size_t(first) + i]);
    dset_raw(g, "Cell ID", H5T_NATIVE_LONG, H5T_NATIVE_LONG, lvec.data(), len);
    for (int d = 0; d < D; ++d) { comp(im.cell_v, d, 1.0); dset_f64(g, std::string("Cell velocity ") + dim_name(d), vec); }
    scal(im.cell_rho); dset_f64(g, "Cell density", vec);
    scal(im.cell_p); dset_f64(g, "Cell pressure", vec);
}

/* Write_Inlet_Data (src/H5IO.cpp:525-541) */
void write_inlet(hid_t g, const std::vector<int64_t>& back, const std::vector<int64_t>& buffer)
{
    std::vector<unsigned long> bk(back.begin(), back.end()), bf(buffer.begin(), buffer.end());
    dset_raw(g, "Back index", H5T_NATIVE_ULONG, H5T_NATIVE_ULONG, bk.data(), bk.size());
    dset_raw(g, "Buffer index", H5T_NATIVE_ULONG, H5T_NATIVE_ULONG, bf.data(), bf.size());
}

/* Write_Zone_Attributes (src/H5IO.cpp:1027-1088) */
void write_zone_attrs(hid_t g, const Image& im, size_t k, const fjsph_restart_meta& mt)
{
    const int D = im.dim;
    const fjsph_block& bl = im.blocks[k];
    const fjsph_block_planes& pl = im.planes[k];
    if (mt.block_names) attr_str(g, "Block name", mt.block_names[k]);
    attr_f64(g, "Insertion normal x", pl.insert_norm[0]);
    attr_f64(g, "Insertion normal y", pl.insert_norm[1]);
    attr_f64(g, "Deletion normal x", pl.delete_norm[0]);
    attr_f64(g, "Deletion normal y", pl.delete_norm[1]);
    attr_f64(g, "Aerodynamic normal x", pl.aero_norm[0]);
    attr_f64(g, "Aerodynamic normal y", pl.aero_norm[1]);
    if (D == 3) {
        attr_f64(g, "Insertion normal z", pl.insert_norm[2]);
        attr_f64(g, "Deletion normal z", pl.delete_norm[2]);
        attr_f64(g, "Aerodynamic normal z", pl.aero_norm[2]);
    }
    attr_f64(g, "Insertion plane constant", pl.insconst);
    attr_f64(g, "Deletion plane constant", pl.delconst);
    attr_f64(g, "Aerodynamic plane constant", pl.aeroconst);
    const size_t nt = im.times[k].size();
    attr_f64(g, "Times count", double(nt));
    if (nt != 0) {
        for (size_t t = 0; t < nt; ++t) {
            const std::string ts = std::to_string(t);
            attr_f64(g, "Time " + ts, im.times[k][t]);
            for (int d = 0; d < D; ++d)
                attr_f64(g, "Velocity " + ts + " " + dim_name(d), im.vels[k][t * 3 + size_t(d)]);
        }
    } else {
        for (int d = 0; d < D; ++d) attr_f64(g, std::string("Velocity ") + dim_name(d), bl.vel[d]);
    }
    const int32_t btype = mt.block_types ? mt.block_types[k] : (pl.inlet ? 6 : 0);
    attr_f64(g, "Fixed velocity or dynamic inlet", double(pl.fixed_vel_or_dynamic));
    attr_f64(g, "Lattice or HCPL packing", double(mt.particle_order ? mt.particle_order[k] : 0));
    attr_f64(g, "Boundary solver type", double(bl.bound_solver));
    attr_f64(g, "Boundary is no slip", double(bl.no_slip));
    attr_f64(g, "Block type", double(btype));
}

void validate(const Image& im)
{
    const int D = im.dim;
    if (D != 2 && D != 3) fail("dim must be 2 or 3", FJSPH_ERR_ARG);
    const size_t n = size_t(im.n);
    auto need = [&](size_t have, size_t want, const char* what) {
        if (have != want) fail(std::string("array size mismatch: ") + what, FJSPH_ERR_ARG);
    };
    need(im.x.size(), n * D, "x"); need(im.v.size(), n * D, "v"); need(im.acc.size(), n * D, "acc");
    need(im.cell_v.size(), n * D, "cell_v");
    for (auto* a : {&im.p, &im.rho, &im.rrho, &im.m, &im.cell_p, &im.cell_rho}) need(a->size(), n, "scalar field");
    need(im.b.size(), n, "b"); need(im.pid.size(), n, "pid"); need(im.cell_id.size(), n, "cell_id");
    int64_t at = 0;
    bool fluid_seen = false;
    for (size_t k = 0; k < im.blocks.size(); ++k) {
        const fjsph_block& bl = im.blocks[k];
        if (bl.first != at || bl.last < bl.first) fail("blocks must tile [0, n) in order", FJSPH_ERR_ARG);
        if (bl.is_fluid) fluid_seen = true;
        else if (fluid_seen) fail("boundary blocks must come first", FJSPH_ERR_ARG);
        at = bl.last;
        if (!im.back[k].empty() && im.buffer[k].size() != im.back[k].size() * FJSPH_RESTART_NBUF)
            fail("the restart layout holds 4 buffer particles per inlet column", FJSPH_ERR_ARG);
    }
    if (at != im.n) fail("blocks must tile [0, n)", FJSPH_ERR_ARG);
}

void write_file(const char* path, const Image& im, const fjsph_restart_meta& mt)
{
    validate(im);
    std::remove(path); /* Write_HDF5 replaces an existing file */
    H5 f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, std::string("creating ") + path);
    int nbb = 0, nfb = 0;
    int64_t bp = 0, fp = 0;
    for (const auto& bl : im.blocks) {
        if (bl.is_fluid) { ++nfb; fp += bl.last - bl.first; }
        else { ++nbb; bp += bl.last - bl.first; }
    }
    write_root(f, im, mt, nbb, nfb, bp, fp);
    int kb = 0, kf = 0;
    for (size_t k = 0; k < im.blocks.size(); ++k) {
        const fjsph_block& bl = im.blocks[k];
        const std::string name = bl.is_fluid ? "Fluid " + std::to_string(kf++) : "Boundary " + std::to_string(kb++);
        H5 g(H5Gcreate2(f, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "group " + name);
        write_zone_data(g, im, mt.scale, bl.first, bl.last);
        if (bl.is_fluid && im.planes[k].inlet && !im.back[k].empty()) write_inlet(g, im.back[k], im.buffer[k]);
        write_zone_attrs(g, im, k, mt);
    }
}

/* ---------------------------------------------------------------- reading */

bool has_attr(hid_t loc, const std::string& name)
{
    return H5Aexists(loc, name.c_str()) > 0;
}

template <class T>
T attr_read(hid_t loc, const std::string& name, hid_t mtype)
{
    if (!has_attr(loc, name)) fail("attribute \"" + name + "\" missing");
    H5 a(H5Aopen(loc, name.c_str(), H5P_DEFAULT), H5Aclose, "opening attribute " + name);
    T v{};
    check(H5Aread(a, mtype, &v), "reading attribute " + name);
    return v;
}

double attr_f64_read(hid_t loc, const std::string& name)
{
    return attr_read<double>(loc, name, H5T_NATIVE_DOUBLE);
}

uint64_t attr_u64_read(hid_t loc, const std::string& name)
{
    return attr_read<unsigned long>(loc, name, H5T_NATIVE_ULONG);
}

uint32_t attr_u32_read(hid_t loc, const std::string& name)
{
    return attr_read<unsigned>(loc, name, H5T_NATIVE_UINT);
}

bool has_dset(hid_t loc, const std::string& name)
{
    return H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0;
}

template <class T>
std::vector<T> dset_read(hid_t loc, const std::string& name, hid_t mtype)
{
    if (!has_dset(loc, name)) fail("dataset \"" + name + "\" missing");
    H5 d(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose, "opening " + name);
    H5 sp(H5Dget_space(d), H5Sclose, "H5Dget_space");
    if (H5Sget_simple_extent_ndims(sp) != 1) fail("dataset \"" + name + "\" is not 1-D");
    hsize_t dims[1] = {0};
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    std::vector<T> v(static_cast<size_t>(dims[0]));
    if (!v.empty()) check(H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "reading " + name);
    return v;
}

size_t dset_len(hid_t loc, const std::string& name)
{
    if (!has_dset(loc, name)) fail("dataset \"" + name + "\" missing");
    H5 d(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose, "opening " + name);
    H5 sp(H5Dget_space(d), H5Sclose, "H5Dget_space");
    hsize_t dims[1] = {0};
    if (H5Sget_simple_extent_ndims(sp) != 1) fail("dataset \"" + name + "\" is not 1-D");
    H5Sget_simple_extent_dims(sp, dims, nullptr);
    return size_t(dims[0]);
}

std::string group_name(int k, int nbb)
{
    return k < nbb ? "Boundary " + std::to_string(k) : "Fluid " + std::to_string(k - nbb);
}

/* Read_HDF5's root attributes (src/H5IO.cpp:1230-1247) and the group sizes. */
fjsph_restart_header read_header(hid_t f)
{
    fjsph_restart_header h{};
    h.clock.current_time = attr_f64_read(f, "Simulation current time");
    h.clock.current_frame = attr_u32_read(f, "Current frame");
    h.particle_step = attr_f64_read(f, "SPH initial spacing");
    h.single_file = int32_t(attr_u32_read(f, "Single file for output"));
    h.frame_time_interval = attr_f64_read(f, "SPH frame time interval");
    h.scale = attr_f64_read(f, "Grid scale");
    h.clock.cfl = attr_f64_read(f, "SPH CFL condition");
    const uint64_t nbb = attr_u64_read(f, "Number of boundary blocks");
    const uint64_t nfb = attr_u64_read(f, "Number of fluid blocks");
    h.clock.next_part_id = int64_t(attr_u64_read(f, "Particle index to add"));
    h.clock.n_stable = attr_u32_read(f, "SPH stable CFL count");
    h.clock.n_unstable = attr_u32_read(f, "SPH unstable CFL count");
    h.clock.last_frame_time = h.clock.current_time;
    h.clock.delta_t = 0.0;
    if (nbb + nfb == 0 || nbb + nfb > 1000000) fail("bad block counts");
    h.n_bound_blocks = int32_t(nbb);
    h.n_fluid_blocks = int32_t(nfb);
    h.dim = 0;
    for (int k = 0; k < int(nbb + nfb); ++k) {
        const std::string gn = group_name(k, int(nbb));
        if (!has_dset(f, gn)) fail("group \"" + gn + "\" missing");
        H5 g(H5Gopen2(f, gn.c_str(), H5P_DEFAULT), H5Gclose, "opening " + gn);
        const int d = has_dset(g, "Position coordinate z") ? 3 : 2;
        if (h.dim == 0) h.dim = d;
        else if (d != h.dim) fail("groups of different dimension");
        const int64_t len = int64_t(dset_len(g, "Position coordinate x"));
        (k < int(nbb) ? h.bound_points : h.fluid_points) += len;
    }
    return h;
}

void read_file(const char* path, Image& im)
{
    H5 f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("opening ") + path);
    const fjsph_restart_header h = read_header(f);
    const int D = h.dim;
    im.dim = D;
    im.clock = h.clock;
    im.particle_step = h.particle_step;
    im.frame_time_interval = h.frame_time_interval;
    im.scale = h.scale;
    im.single_file = h.single_file;
    im.max_points = has_attr(f, "SPH maximum particle count") ? int64_t(attr_u64_read(f, "SPH maximum particle count"))
                                                              : 9999999;
    const int nbk = h.n_bound_blocks + h.n_fluid_blocks;
    im.blocks.assign(size_t(nbk), fjsph_block{});
    im.planes.assign(size_t(nbk), fjsph_block_planes{});
    im.times.assign(size_t(nbk), {});
    im.vels.assign(size_t(nbk), {});
    im.back.assign(size_t(nbk), {});
    im.buffer.assign(size_t(nbk), {});
    im.n = h.bound_points + h.fluid_points;
    im.nb = h.bound_points;
    const size_t n = size_t(im.n);
    im.x.assign(n * D, 0.0); im.v.assign(n * D, 0.0); im.acc.assign(n * D, 0.0); im.cell_v.assign(n * D, 0.0);
    for (auto* a : {&im.p, &im.rho, &im.rrho, &im.m, &im.cell_p, &im.cell_rho}) a->assign(n, 0.0);
    im.b.assign(n, 0);
    im.pid.assign(n, 0);
    im.cell_id.assign(n, 0);
    int64_t at = 0;
    for (int k = 0; k < nbk; ++k) {
        const std::string gn = group_name(k, h.n_bound_blocks);
        H5 g(H5Gopen2(f, gn.c_str(), H5P_DEFAULT), H5Gclose, "opening " + gn);
        const size_t len = dset_len(g, "Position coordinate x");
        fjsph_block& bl = im.blocks[size_t(k)];
        bl.first = at;
        bl.last = at + int64_t(len);
        bl.is_fluid = k >= h.n_bound_blocks ? 1 : 0;
        /* Read_Zone_Data (src/H5IO.cpp:801-898) */
        auto vecf = [&](const char* base, std::vector<double>& dst, double mul) {
            for (int d = 0; d < D; ++d) {
                const std::vector<double> c = dset_read<double>(g, std::string(base) + " " + dim_name(d), H5T_NATIVE_DOUBLE);
                if (c.size() != len) fail(gn + ": dataset lengths differ");
                for (size_t i = 0; i < len; ++i) dst[(size_t(at) + i) * D + d] = c[i] * mul;
            }
        };
        auto scal = [&](const char* name, std::vector<double>& dst) {
            const std::vector<double> c = dset_read<double>(g, name, H5T_NATIVE_DOUBLE);
            if (c.size() != len) fail(gn + ": dataset lengths differ");
            std::copy(c.begin(), c.end(), dst.begin() + at);
        };
        vecf("Position coordinate", im.x, h.scale);
        vecf("Velocity", im.v, 1.0);
        vecf("Acceleration", im.acc, 1.0);
        scal("Pressure", im.p);
        scal("Density", im.rho);
        scal("Density gradient", im.rrho);
        scal("Mass", im.m);
        const std::vector<int> bc = dset_read<int>(g, "Boundary condition", H5T_NATIVE_INT);
        const std::vector<unsigned long> ids = dset_read<unsigned long>(g, "Particle ID", H5T_NATIVE_ULONG);
        const std::vector<long> cid = dset_read<long>(g, "Cell ID", H5T_NATIVE_LONG);
        if (bc.size() != len || ids.size() != len || cid.size() != len) fail(gn + ": dataset lengths differ");
        for (size_t i = 0; i < len; ++i) {
            im.b[size_t(at) + i] = bc[i];
            im.pid[size_t(at) + i] = int64_t(ids[i]);
            im.cell_id[size_t(at) + i] = int64_t(cid[i]);
        }
        scal("Cell density", im.cell_rho);
        scal("Cell pressure", im.cell_p);
        vecf("Cell velocity", im.cell_v, 1.0);
        /* block attributes (what Write_Zone_Attributes wrote) */
        fjsph_block_planes& pl = im.planes[size_t(k)];
        for (int d = 0; d < 3; ++d) pl.insert_norm[d] = pl.delete_norm[d] = pl.aero_norm[d] = 0.0;
        for (int d = 0; d < D; ++d) {
            pl.insert_norm[d] = attr_f64_read(g, std::string("Insertion normal ") + dim_name(d));
            pl.delete_norm[d] = attr_f64_read(g, std::string("Deletion normal ") + dim_name(d));
            pl.aero_norm[d] = attr_f64_read(g, std::string("Aerodynamic normal ") + dim_name(d));
        }
        pl.insconst = attr_f64_read(g, "Insertion plane constant");
        pl.delconst = attr_f64_read(g, "Deletion plane constant");
        pl.aeroconst = attr_f64_read(g, "Aerodynamic plane constant");
        pl.fixed_vel_or_dynamic = int32_t(attr_f64_read(g, "Fixed velocity or dynamic inlet"));
        pl.inlet = attr_f64_read(g, "Block type") == 6.0 ? 1 : 0;
        bl.bound_solver = int32_t(attr_f64_read(g, "Boundary solver type"));
        bl.no_slip = int32_t(attr_f64_read(g, "Boundary is no slip"));
        const double ntd = attr_f64_read(g, "Times count");
        if (!(ntd >= 0.0) || ntd > 1e7 || ntd != std::floor(ntd)) fail(gn + ": bad \"Times count\"");
        const size_t nt = size_t(ntd);
        bl.n_times = int32_t(nt);
        for (int d = 0; d < 3; ++d) bl.vel[d] = 0.0;
        if (nt > 0) {
            im.times[size_t(k)].resize(nt);
            im.vels[size_t(k)].assign(nt * 3, 0.0);
            for (size_t t = 0; t < nt; ++t) {
                const std::string ts = std::to_string(t);
                im.times[size_t(k)][t] = attr_f64_read(g, "Time " + ts);
                for (int d = 0; d < D; ++d)
                    im.vels[size_t(k)][t * 3 + size_t(d)] = attr_f64_read(g, "Velocity " + ts + " " + dim_name(d));
            }
        } else {
            for (int d = 0; d < D; ++d) bl.vel[d] = attr_f64_read(g, std::string("Velocity ") + dim_name(d));
        }
        /* Read_Inlet_Data (src/H5IO.cpp:900-911) */
        if (bl.is_fluid && has_dset(g, "Back index")) {
            const std::vector<unsigned long> bk = dset_read<unsigned long>(g, "Back index", H5T_NATIVE_ULONG);
            const std::vector<unsigned long> bf = dset_read<unsigned long>(g, "Buffer index", H5T_NATIVE_ULONG);
            if (bf.size() != bk.size() * FJSPH_RESTART_NBUF) fail(gn + ": \"Buffer index\" is not 4 per column");
            im.back[size_t(k)].assign(bk.begin(), bk.end());
            im.buffer[size_t(k)].assign(bf.begin(), bf.end());
        }
        at += int64_t(len);
    }
}

/* ------------------------------------------------------------ host image */

Image image_of_host(const fjsph_restart_host* st)
{
    Image im;
    im.dim = st->dim;
    im.n = st->n;
    im.nb = st->n_bound;
    im.s = st->settings;
    im.clock = st->clock;
    im.max_points = st->max_points;
    if (st->n <= 0 || st->n_blocks <= 0 || !st->blocks || !st->planes) fail("empty state", FJSPH_ERR_ARG);
    const size_t nbk = size_t(st->n_blocks), n = size_t(st->n), D = size_t(st->dim);
    im.blocks.assign(st->blocks, st->blocks + nbk);
    im.planes.assign(st->planes, st->planes + nbk);
    im.times.assign(nbk, {});
    im.vels.assign(nbk, {});
    im.back.assign(nbk, {});
    im.buffer.assign(nbk, {});
    for (size_t k = 0; k < nbk; ++k) {
        const int32_t nt = st->n_times ? st->n_times[k] : 0;
        if (nt > 0) {
            if (!st->times || !st->vels || !st->times[k] || !st->vels[k]) fail("motion table missing", FJSPH_ERR_ARG);
            im.times[k].assign(st->times[k], st->times[k] + nt);
            im.vels[k].assign(st->vels[k], st->vels[k] + size_t(nt) * 3);
        }
        const int32_t nc = st->n_cols ? st->n_cols[k] : 0;
        if (nc > 0) {
            if (!st->back || !st->buffer || !st->back[k] || !st->buffer[k]) fail("inlet columns missing", FJSPH_ERR_ARG);
            im.back[k].assign(st->back[k], st->back[k] + nc);
            im.buffer[k].assign(st->buffer[k], st->buffer[k] + size_t(nc) * FJSPH_RESTART_NBUF);
        }
    }
    auto take = [&](const double* a, size_t w, std::vector<double>& dst) {
        if (!a) fail("field array missing", FJSPH_ERR_ARG);
        dst.assign(a, a + n * w);
    };
    take(st->x, D, im.x); take(st->v, D, im.v); take(st->acc, D, im.acc); take(st->cell_v, D, im.cell_v);
    take(st->p, 1, im.p); take(st->rho, 1, im.rho); take(st->rrho, 1, im.rrho); take(st->m, 1, im.m);
    take(st->cell_p, 1, im.cell_p); take(st->cell_rho, 1, im.cell_rho);
    if (!st->b || !st->pid || !st->cell_id) fail("field array missing", FJSPH_ERR_ARG);
    im.b.assign(st->b, st->b + n);
    im.pid.assign(st->pid, st->pid + n);
    im.cell_id.assign(st->cell_id, st->cell_id + n);
    return im;
}

/* ---------------------------------------------------------- engine image */

void eng(fjsph_ctx* ctx, fjsph_status e, const char* what)
{
    if (e != FJSPH_OK) fail(std::string(what) + ": " + fjsph_last_error(ctx), e);
}

Image image_of_ctx(fjsph_ctx* ctx)
{
    Image im;
    eng(ctx, fjsph_get_settings(ctx, &im.s), "get_settings");
    eng(ctx, fjsph_get_clock(ctx, &im.clock), "get_clock");
    eng(ctx, fjsph_get_counts(ctx, &im.n, &im.nb), "get_counts");
    eng(ctx, fjsph_get_max_points(ctx, &im.max_points), "get_max_points");
    im.dim = im.s.dim;
    const int D = im.dim;
    int32_t nbk = 0;
    eng(ctx, fjsph_get_blocks(ctx, &nbk, nullptr, 0), "get_blocks");
    im.blocks.resize(size_t(nbk));
    eng(ctx, fjsph_get_blocks(ctx, &nbk, im.blocks.data(), nbk), "get_blocks");
    im.planes.resize(size_t(nbk));
    im.times.assign(size_t(nbk), {});
    im.vels.assign(size_t(nbk), {});
    im.back.assign(size_t(nbk), {});
    im.buffer.assign(size_t(nbk), {});
    for (int32_t k = 0; k < nbk; ++k) {
        eng(ctx, fjsph_get_block_planes(ctx, k, &im.planes[size_t(k)]), "get_block_planes");
        int32_t nt = 0;
        eng(ctx, fjsph_get_block_motion(ctx, k, &nt, nullptr, nullptr), "get_block_motion");
        if (nt > 0) {
            im.times[size_t(k)].resize(size_t(nt));
            im.vels[size_t(k)].resize(size_t(nt) * 3);
            eng(ctx, fjsph_get_block_motion(ctx, k, &nt, im.times[size_t(k)].data(), im.vels[size_t(k)].data()),
                "get_block_motion");
        }
        int32_t nc = 0, nb = 0;
        eng(ctx, fjsph_get_inlet(ctx, k, &nc, &nb, nullptr, nullptr), "get_inlet");
        if (nc > 0) {
            if (nb != FJSPH_RESTART_NBUF)
                fail("the restart layout holds 4 buffer particles per inlet column; block " + std::to_string(k) +
                         " has " + std::to_string(nb),
                     FJSPH_ERR_UNSUPPORTED);
            im.back[size_t(k)].resize(size_t(nc));
            im.buffer[size_t(k)].resize(size_t(nc) * FJSPH_RESTART_NBUF);
            eng(ctx, fjsph_get_inlet(ctx, k, &nc, &nb, im.back[size_t(k)].data(), im.buffer[size_t(k)].data()),
                "get_inlet");
        }
    }
    const size_t n = size_t(im.n);
    auto field = [&](int f, size_t w, std::vector<double>& dst) {
        dst.resize(n * w);
        eng(ctx, fjsph_download_field(ctx, FJSPH_STATE_NP1, f, dst.data()), "download_field");
    };
    field(FJSPH_F_X, D, im.x); field(FJSPH_F_V, D, im.v); field(FJSPH_F_ACC, D, im.acc);
    field(FJSPH_F_CELLV, D, im.cell_v);
    field(FJSPH_F_P, 1, im.p); field(FJSPH_F_RHO, 1, im.rho); field(FJSPH_F_RRHO, 1, im.rrho);
    field(FJSPH_F_M, 1, im.m); field(FJSPH_F_CELLP, 1, im.cell_p); field(FJSPH_F_CELLRHO, 1, im.cell_rho);
    std::vector<double> t;
    field(FJSPH_F_B, 1, t);
    im.b.resize(n);
    for (size_t i = 0; i < n; ++i) im.b[i] = int32_t(t[i]);
    field(FJSPH_F_PID, 1, t);
    im.pid.resize(n);
    for (size_t i = 0; i < n; ++i) im.pid[i] = int64_t(t[i]);
    field(FJSPH_F_CELLID, 1, t);
    im.cell_id.resize(n);
    for (size_t i = 0; i < n; ++i) im.cell_id[i] = int64_t(t[i]);
    return im;
}

void image_to_ctx(fjsph_ctx* ctx, const Image& im, int32_t n_blocks, const fjsph_block* blocks,
                  const fjsph_block_planes* planes)
{
    fjsph_settings s{};
    eng(ctx, fjsph_get_settings(ctx, &s), "get_settings");
    if (s.dim != im.dim) fail("the file is " + std::to_string(im.dim) + "-D, the context " + std::to_string(s.dim) + "-D");
    if (std::fabs(s.particle_step - im.particle_step) > 1e-12 * std::fabs(im.particle_step))
        fail("the context's particle_step differs from the file's \"SPH initial spacing\"");
    const size_t nbk = im.blocks.size();
    if (blocks && size_t(n_blocks) != nbk)
        fail("the file holds " + std::to_string(nbk) + " blocks, " + std::to_string(n_blocks) + " given", FJSPH_ERR_ARG);
    /* Init_Particles_Restart: block definitions; Read_HDF5: their ranges */
    std::vector<fjsph_block> bl(im.blocks);
    for (size_t k = 0; k < nbk; ++k) {
        if (blocks) {
            if (blocks[k].is_fluid != im.blocks[k].is_fluid) fail("block kinds differ from the file's", FJSPH_ERR_ARG);
            bl[k] = blocks[k];
            bl[k].first = im.blocks[k].first;
            bl[k].last = im.blocks[k].last;
            bl[k].n_times = int32_t(im.times[k].size());
        }
    }
    eng(ctx, fjsph_set_blocks(ctx, int32_t(nbk), bl.data()), "set_blocks");
    for (size_t k = 0; k < nbk; ++k)
        if (!im.times[k].empty())
            eng(ctx, fjsph_set_block_motion(ctx, int32_t(k), int32_t(im.times[k].size()), im.times[k].data(),
                                            im.vels[k].data()),
                "set_block_motion");
    eng(ctx, fjsph_upload_state(ctx, im.n, im.nb, im.x.data(), im.v.data(), im.rho.data(), im.p.data(), im.m.data(),
                                im.b.data()),
        "upload_state");
    const size_t n = size_t(im.n);
    std::vector<double> t(n);
    for (int which : {FJSPH_STATE_N, FJSPH_STATE_NP1}) {
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_ACC, im.acc.data()), "upload ACC");
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_RRHO, im.rrho.data()), "upload RRHO");
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_CELLV, im.cell_v.data()), "upload CELLV");
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_CELLP, im.cell_p.data()), "upload CELLP");
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_CELLRHO, im.cell_rho.data()), "upload CELLRHO");
        for (size_t i = 0; i < n; ++i) t[i] = double(im.pid[i]);
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_PID, t.data()), "upload PID");
        for (size_t i = 0; i < n; ++i) t[i] = double(im.cell_id[i]);
        eng(ctx, fjsph_upload_field(ctx, which, FJSPH_F_CELLID, t.data()), "upload CELLID");
    }
    for (size_t k = 0; k < nbk; ++k) {
        const fjsph_block_planes& pl = planes ? planes[k] : im.planes[k];
        if (bl[k].is_fluid) eng(ctx, fjsph_set_block_planes(ctx, int32_t(k), &pl), "set_block_planes");
        if (bl[k].is_fluid && pl.inlet && !im.back[k].empty())
            eng(ctx, fjsph_set_inlet(ctx, int32_t(k), int32_t(im.back[k].size()), FJSPH_RESTART_NBUF,
                                     im.back[k].data(), im.buffer[k].data()),
                "set_inlet");
    }
    fjsph_clock c{};
    eng(ctx, fjsph_get_clock(ctx, &c), "get_clock");
    c.current_time = im.clock.current_time;
    c.last_frame_time = im.clock.current_time; /* src/H5IO.cpp:1248 */
    c.cfl = im.clock.cfl;
    c.current_frame = im.clock.current_frame;
    c.n_stable = im.clock.n_stable;
    c.n_unstable = im.clock.n_unstable;
    c.next_part_id = im.clock.next_part_id;
    eng(ctx, fjsph_set_clock(ctx, &c), "set_clock");
}

template <class F>
fjsph_status guarded(F&& f)
{
    try {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
        f();
        return FJSPH_OK;
    } catch (const RestartError& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return FJSPH_ERR_STATE;
    }
}

}  // namespace

extern "C" {

int32_t fjsph_restart_abi_version(void)
{
    return FJSPH_RESTART_ABI_VERSION;
}

const char* fjsph_restart_last_error(void)
{
    return g_err.c_str();
}

void fjsph_restart_meta_default(fjsph_restart_meta* m, int32_t dim)
{
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->scale = 1.0;
    m->hydro_height = -1.0;
    m->bound_step_factor = 1.0;
    m->restart_tol = 0.001;
    m->contangb = 150.0;
    m->v_ref = 0.0;
    m->M_ref = -1.0;
    m->max_x_sph = 9999999.0;
    m->max_x = 9999999.0;
    m->offset_axis = dim == 2 ? 2 : 0;
    m->single_file = 1;
    m->out_encoding = 1;
    m->max_frames = -1; /* uint max_frames = -1 (src/Var.h:70) */
    m->ipt_eq_order = 2;
    m->streak_out = 1;
    m->foam_label_size = 32;
    m->foam_scalar_size = 64;
}

fjsph_status fjsph_restart_read_header(const char* path, fjsph_restart_header* out)
{
    if (!path || !out) return FJSPH_ERR_ARG;
    return guarded([&] {
        H5 f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, std::string("opening ") + path);
        *out = read_header(f);
    });
}

fjsph_status fjsph_restart_write_host(const char* path, const fjsph_restart_host* st, const fjsph_restart_meta* meta)
{
    if (!path || !st) return FJSPH_ERR_ARG;
    return guarded([&] {
        fjsph_restart_meta m;
        fjsph_restart_meta_default(&m, st->dim);
        write_file(path, image_of_host(st), meta ? *meta : m);
    });
}

fjsph_status fjsph_restart_read_host(const char* path, fjsph_restart_host* st)
{
    if (!path || !st || !st->blocks) return FJSPH_ERR_ARG;
    return guarded([&] {
        Image im;
        read_file(path, im);
        const size_t nbk = im.blocks.size();
        if (size_t(st->n_blocks) < nbk) fail("n_blocks smaller than the file's block count", FJSPH_ERR_ARG);
        st->n_blocks = int32_t(nbk);
        st->dim = im.dim;
        st->n = im.n;
        st->n_bound = im.nb;
        st->clock = im.clock;
        st->max_points = im.max_points;
        st->settings.particle_step = im.particle_step;
        st->settings.frame_time_interval = im.frame_time_interval;
        for (size_t k = 0; k < nbk; ++k) {
            st->blocks[k] = im.blocks[k];
            if (st->planes) st->planes[k] = im.planes[k];
            if (st->n_times) st->n_times[k] = int32_t(im.times[k].size());
            if (st->n_cols) st->n_cols[k] = int32_t(im.back[k].size());
        }
        if (!st->x) return; /* first call: sizes only */
        const size_t n = size_t(im.n), D = size_t(im.dim);
        auto put = [&](const std::vector<double>& a, double* dst, size_t w) {
            if (!dst) fail("field array missing", FJSPH_ERR_ARG);
            std::copy(a.begin(), a.begin() + n * w, dst);
        };
        put(im.x, st->x, D); put(im.v, st->v, D); put(im.acc, st->acc, D); put(im.cell_v, st->cell_v, D);
        put(im.p, st->p, 1); put(im.rho, st->rho, 1); put(im.rrho, st->rrho, 1); put(im.m, st->m, 1);
        put(im.cell_p, st->cell_p, 1); put(im.cell_rho, st->cell_rho, 1);
        if (!st->b || !st->pid || !st->cell_id) fail("field array missing", FJSPH_ERR_ARG);
        std::copy(im.b.begin(), im.b.end(), st->b);
        std::copy(im.pid.begin(), im.pid.end(), st->pid);
        std::copy(im.cell_id.begin(), im.cell_id.end(), st->cell_id);
        for (size_t k = 0; k < nbk; ++k) {
            if (!im.times[k].empty()) {
                if (!st->times || !st->vels || !st->times[k] || !st->vels[k]) fail("motion table missing", FJSPH_ERR_ARG);
                std::copy(im.times[k].begin(), im.times[k].end(), st->times[k]);
                std::copy(im.vels[k].begin(), im.vels[k].end(), st->vels[k]);
            }
            if (!im.back[k].empty()) {
                if (!st->back || !st->buffer || !st->back[k] || !st->buffer[k]) fail("inlet columns missing", FJSPH_ERR_ARG);
                std::copy(im.back[k].begin(), im.back[k].end(), st->back[k]);
                std::copy(im.buffer[k].begin(), im.buffer[k].end(), st->buffer[k]);
            }
        }
    });
}

fjsph_status fjsph_restart_write(fjsph_ctx* ctx, const char* path, const fjsph_restart_meta* meta)
{
    if (!ctx || !path) return FJSPH_ERR_ARG;
    return guarded([&] {
        const Image im = image_of_ctx(ctx);
        fjsph_restart_meta m;
        fjsph_restart_meta_default(&m, im.dim);
        write_file(path, im, meta ? *meta : m);
    });
}

fjsph_status fjsph_restart_read(fjsph_ctx* ctx, const char* path, int32_t n_blocks, const fjsph_block* blocks,
                                const fjsph_block_planes* planes)
{
    if (!ctx || !path || (blocks && n_blocks <= 0)) return FJSPH_ERR_ARG;
    return guarded([&] {
        Image im;
        read_file(path, im);
        image_to_ctx(ctx, im, n_blocks, blocks, planes);
    });
}

}  // extern "C"
