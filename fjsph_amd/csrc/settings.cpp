/*
 * Host-side settings of the engine: reference defaults, the derived
 * constants of Set_Values, and a reader for the reference's para files.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/fjsph_engine.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

/* get_n_full (src/Geometry.cpp:282-308): lattice points stepped by repeated
 * addition of dx from -2(H+dx) to 2(H+dx), counted strictly inside 4H^2. */
double lattice_full_count(int dim, double dx, double H)
{
    const double lim = 2.0 * (H + dx);
    const double sr = 4.0 * H * H;
    long n = 0;
    for (double x = -lim; x <= lim; x += dx)
        for (double y = -lim; y <= lim; y += dx) {
            const double xy = x * x + y * y;
            if (dim == 2) {
                n += xy < sr;
                continue;
            }
            for (double z = -lim; z <= lim; z += dx) n += (xy + z * z) < sr;
        }
    return double(n);
}

std::string trim(const std::string& s)
{
    const char* ws = " \n\r\t\f\v";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

bool read_vec(const std::string& v, double* out, int n)
{
    std::string t = v;
    for (char& ch : t)
        if (ch == ',') ch = ' ';
    std::istringstream is(t);
    int k = 0;
    double d;
    while (k < n && is >> d) out[k++] = d;
    return k > 0;
}

}  // namespace

extern "C" void fjsph_settings_default(fjsph_settings* s, int32_t dim)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->particle_step = -1.0;
    s->H_fac = 2.0;
    s->rho_rest = 1000.0;
    s->speed_sound = 300.0;
    s->gam = 7.0;
    s->mu = 8.94e-4;
    s->sig = 0.0708;
    s->visc_alpha = 0.1;
    s->dsph_delta = 0.1;
    s->press_back = 0.0;
    s->press_pipe = 0.0;
    s->rho_var = 50.0;
    s->rho_max = 1500.0;
    s->rho_min = 500.0;
    s->rho_max_iter = 1.0;
    s->grav[0] = 0.0;
    s->grav[1] = dim == 3 ? 0.0 : -9.81;
    s->grav[2] = dim == 3 ? -9.81 : 0.0;
    s->cfl = 1.0;
    s->cfl_min = 0.1;
    s->cfl_max = 2.0;
    s->cfl_step = 0.05;
    s->subits_factor = 0.333;
    s->min_residual = -7.0;
    s->delta_t_max = 1.0;
    s->delta_t_min = 0.0;
    s->frame_time_interval = 1.0e30;
    s->max_shift_vel = 9999999.0;
    s->p_ref = 101353.0;
    s->rho_g = 1.29251;
    s->mu_g = 1.716e-5;
    s->temp_g = 298.0;
    s->R_g = 287.0;
    s->gamma_g = 1.403;
    s->lam_cutoff = 0.75;
    s->i_interp_fac = 0.5;
    s->Cf = 1.0 / 3.0;
    s->Ck = 8.0;
    s->Cd = 5.0;
    s->Cb = 0.5;
    s->dim = dim;
    s->precision = FJSPH_PREC_FP64;
    s->integrator = FJSPH_NEWMARK_BETA;
    s->eos = FJSPH_COLE_EOS;
    s->aero_case = FJSPH_NO_AERO;
    s->aero_source = FJSPH_CONST_VEL;
    s->use_lam = 1;
    s->use_TAB_def = 0;
    s->ale = 0;
    s->max_subits = 20;
    s->n_stable_limit = 10;
    s->n_unstable_limit = 3;
}

extern "C" fjsph_status fjsph_derive(const fjsph_settings* s, fjsph_derived* o)
{
    if (!s || !o) return FJSPH_ERR_ARG;
    if (s->dim != 2 && s->dim != 3) return FJSPH_ERR_ARG;
    if (!(s->particle_step > 0.0) || !(s->H_fac > 0.0)) return FJSPH_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    const int D = s->dim;
    const double ps = s->particle_step;
    o->B = s->rho_rest * s->speed_sound * s->speed_sound / s->gam;
    const double rho_pipe = (s->eos == FJSPH_COLE_EOS)
                                ? s->rho_rest * std::pow((s->press_pipe - s->press_back) / o->B + 1.0, 1.0 / s->gam)
                                : (s->press_pipe - s->press_back) / (s->speed_sound * s->speed_sound) + s->rho_rest;
    const bool unset_limits = (s->rho_max == 1500.0 && s->rho_min == 500.0);
    o->rho_max = unset_limits ? s->rho_rest * (1.0 + 0.01 * s->rho_var) : s->rho_max;
    o->rho_min = unset_limits ? s->rho_rest * (1.0 - 0.01 * s->rho_var) : s->rho_min;
    o->dx = ps * std::pow(rho_pipe / s->rho_rest, 1.0 / D);
    o->sim_mass = s->rho_rest * std::pow(ps, D);
    o->bnd_mass = o->sim_mass;
    o->sos = std::sqrt(s->temp_g * s->R_g * s->gamma_g);
    o->i_sos_sq = 1.0 / (o->sos * o->sos);
    o->H = s->H_fac * ps;
    o->H_sq = o->H * o->H;
    o->sr = 4.0 * o->H_sq;
    o->dsph_cont = 2.0 * s->dsph_delta * o->H * s->speed_sound;
    o->nu = s->mu / s->rho_rest;
    o->W_correc = (D == 2) ? 7.0 / (4.0 * kPi * o->H_sq) : 21.0 / (16.0 * kPi * o->H_sq * o->H);
    const double t = 1.0 - 0.5 * ps / o->H;
    o->W_dx = t * t * t * t * (2.0 * ps / o->H + 1.0) * o->W_correc;
    /* AERO::GetYcoef with diam = particle_step (src/Var.h:244-266) */
    if (D == 3) {
        o->L = ps * std::cbrt(3.0 / (4.0 * kPi));
        o->A_sphere = kPi * o->L * o->L;
    } else {
        o->L = ps / std::sqrt(kPi);
        o->A_sphere = 2.0 * o->L;
    }
    o->td = 2.0 * s->rho_rest * std::pow(o->L, D - 1) / (s->Cd * s->mu);
    o->omega = std::sqrt(s->Ck * s->sig / (s->rho_rest * std::pow(o->L, D)) - 1.0 / (o->td * o->td));
    o->tmax = -2.0 * (std::atan(std::sqrt((o->td * o->omega) * (o->td * o->omega) + 1.0) + o->td * o->omega) - kPi) / o->omega;
    o->Cdef = 1.0 - std::exp(-o->tmax / o->td) *
                        (std::cos(o->omega * o->tmax) + std::sin(o->omega * o->tmax) / (o->omega * o->td));
    o->ycoef = 0.5 * o->Cdef * (s->Cf / (s->Ck * s->Cb)) * (s->rho_g * o->L) / s->sig;
    o->n_full = lattice_full_count(D, ps, o->H);
    o->i_n_full = 1.0 / o->n_full;
    o->interp_fac = 1.0 / s->i_interp_fac;
    o->A_plate = (D == 3) ? ps * ps : ps;
    /* pairwise surface-tension normaliser (src/Resid.cpp:437-439) */
    const double h2 = 2.0 * o->H;
    o->lam_st = 6.0 / 81.0 * h2 * h2 * h2 / (kPi * kPi * kPi * kPi) * (9.0 / 4.0 * kPi * kPi * kPi - 6.0 * kPi - 4.0);
    return FJSPH_OK;
}

extern "C" fjsph_status fjsph_settings_parse_file(fjsph_settings* s, const char* path)
{
    if (!s || !path) return FJSPH_ERR_ARG;
    std::ifstream in(path);
    if (!in) return FJSPH_ERR_ARG;
    std::string line;
    int bad = 0;
    while (std::getline(in, line)) {
        line = trim(line);
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        const size_t colon = line.find(':');
        if (colon == std::string::npos) continue;
        const std::string key = trim(line.substr(0, colon));
        const std::string val = trim(line.substr(colon + 1));
        double d = 0.0;
        const bool num = (std::istringstream(val) >> d) ? true : false;
        auto setd = [&](const char* k, double& dst) { if (key == k && num) dst = d; };
        auto seti = [&](const char* k, int32_t& dst) { if (key == k && num) dst = int32_t(d); };
        setd("SPH initial spacing", s->particle_step);
        setd("SPH smoothing length factor", s->H_fac);
        setd("Reference dispersed density", s->rho_rest);
        setd("SPH speed of sound", s->speed_sound);
        setd("Reference dispersed viscosity", s->mu);
        setd("Reference surface tension", s->sig);
        setd("SPH artificial viscosity factor", s->visc_alpha);
        setd("SPH delta coefficient", s->dsph_delta);
        setd("SPH background pressure", s->press_back);
        setd("SPH starting pressure", s->press_pipe);
        setd("SPH maximum absolute density variation (%)", s->rho_var);
        setd("SPH density variation to reduce timestep (%)", s->rho_max_iter);
        setd("SPH maximum density", s->rho_max);
        setd("SPH minimum density", s->rho_min);
        setd("SPH CFL condition", s->cfl);
        setd("SPH minimum CFL", s->cfl_min);
        setd("SPH maximum CFL", s->cfl_max);
        setd("SPH unstable CFL step", s->cfl_step);
        setd("SPH stable CFL count iteration factor", s->subits_factor);
        setd("SPH solver minimum residual", s->min_residual);
        setd("SPH maximum timestep", s->delta_t_max);
        setd("SPH minimum timestep", s->delta_t_min);
        setd("SPH frame time interval", s->frame_time_interval);
        setd("SPH maximum shifting velocity", s->max_shift_vel);
        setd("Reference pressure", s->p_ref);
        setd("Reference density", s->rho_g);
        setd("Sutherland reference viscosity", s->mu_g);
        setd("Reference temperature", s->temp_g);
        setd("Gas constant gamma", s->gamma_g);
        setd("SPH aerodynamic cutoff value", s->lam_cutoff);
        setd("SPH aerodynamic interpolation factor", s->i_interp_fac);
        seti("SPH equation of state (0=Cole/1=Isothermal)", s->eos);
        seti("SPH interpolation factor (0=ncount/1=lambda)", s->use_lam);
        seti("SPH use TAB deformation (0/1)", s->use_TAB_def);
        seti("SPH Newmark-Beta iteration limit", s->max_subits);
        seti("SPH stable CFL count limit", s->n_stable_limit);
        seti("SPH unstable CFL count limit", s->n_unstable_limit);
        if (key == "SPH gravity vector") {
            double g[3] = {0.0, 0.0, 0.0};
            if (read_vec(val, g, s->dim)) for (int k = 0; k < 3; ++k) s->grav[k] = k < s->dim ? g[k] : 0.0;
        }
        if (key == "SPH freestream velocity") {
            double g[3] = {0.0, 0.0, 0.0};
            if (read_vec(val, g, s->dim)) for (int k = 0; k < 3; ++k) s->v_inf[k] = k < s->dim ? g[k] : 0.0;
        }
        if (key == "SPH integration solver") {
            if (val == "Newmark-Beta") s->integrator = FJSPH_NEWMARK_BETA;
            else if (val == "Runge-Kutta") s->integrator = FJSPH_RUNGE_KUTTA;
            else bad = 1;
        }
        if (key == "SPH aerodynamic case") {
            if (val == "(none)") s->aero_case = FJSPH_NO_AERO;
            else if (val == "Gissler") s->aero_case = FJSPH_GISSLER;
            else if (val == "Induced_pressure") s->aero_case = FJSPH_INDUCED_PRESSURE;
            else if (val == "Skin_friction") s->aero_case = FJSPH_SKIN_FRIC;
            else bad = 1;
        }
    }
    return bad ? FJSPH_ERR_ARG : FJSPH_OK;
}
