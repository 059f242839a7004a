"""ctypes mirror of include/fjsph_engine.h (structs, enums, defaults).

The struct layouts must match the header field for field; tests check the
sizes against the engine library and the defaults against
fjsph_settings_default().
"""
import ctypes as C

import numpy as np

ABI_VERSION = 7

# fjsph_status
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_STATE, ERR_UNSUPPORTED, ERR_NO_DEVICE = range(7)

# particle types (src/VarDefs.h:92-102)
BOUND, PISTON, BUFFER, BACK, PIPE, FREE, OUTLET, LOST = range(8)

PREC_FP64, PREC_MIXED = 0, 1
NEWMARK_BETA, RUNGE_KUTTA = 0, 1
DBC, PRESSURE_G, GHOST = 0, 1, 2
COLE_EOS, ISO_EOS = 0, 1
NO_AERO, GISSLER, INDUCED_PRESSURE, SKIN_FRIC = 0, 1, 2, 3
CONST_VEL, MESH_INFL, VLM_INFL, SURROGATE = 0, 1, 2, 3
STATE_N, STATE_NP1 = 0, 1

FIELDS = [
    "X", "V", "ACC", "AF", "AVISC", "RHO", "P", "RRHO", "M", "DELTAD", "L",
    "GRADRHO", "NORM", "LAM", "LAM_NB", "COLOURG", "COLOUR", "KERNSUM", "SURF",
    "CURVE", "NORM_CURVE", "WOCCL", "PDIST", "CELLV", "CELLP", "CELLRHO",
    "CELLID", "VPERT", "SURFZONE", "B", "PID", "NFAILED", "INTERNAL",
]
F = {name: i for i, name in enumerate(FIELDS)}
VECTOR_FIELDS = {"X", "V", "ACC", "AF", "AVISC", "GRADRHO", "NORM", "CELLV", "VPERT"}


def field_width(name, dim):
    if name in VECTOR_FIELDS:
        return dim
    if name == "L":
        return dim * dim
    return 1


TIMER_NAMES = ["neighbour", "dsph", "surface", "dissipation", "shift",
               "boundary", "force", "reduce", "update", "aero"]


class Settings(C.Structure):
    _fields_ = [
        ("particle_step", C.c_double), ("H_fac", C.c_double), ("rho_rest", C.c_double),
        ("speed_sound", C.c_double), ("gam", C.c_double), ("mu", C.c_double),
        ("sig", C.c_double), ("visc_alpha", C.c_double), ("dsph_delta", C.c_double),
        ("press_back", C.c_double), ("press_pipe", C.c_double), ("rho_var", C.c_double),
        ("rho_max", C.c_double), ("rho_min", C.c_double), ("rho_max_iter", C.c_double),
        ("grav", C.c_double * 3),
        ("cfl", C.c_double), ("cfl_min", C.c_double), ("cfl_max", C.c_double),
        ("cfl_step", C.c_double), ("subits_factor", C.c_double), ("min_residual", C.c_double),
        ("delta_t_max", C.c_double), ("delta_t_min", C.c_double),
        ("frame_time_interval", C.c_double), ("max_shift_vel", C.c_double),
        ("v_inf", C.c_double * 3),
        ("p_ref", C.c_double), ("rho_g", C.c_double), ("mu_g", C.c_double),
        ("temp_g", C.c_double), ("R_g", C.c_double), ("gamma_g", C.c_double),
        ("lam_cutoff", C.c_double), ("i_interp_fac", C.c_double),
        ("Cf", C.c_double), ("Ck", C.c_double), ("Cd", C.c_double), ("Cb", C.c_double),
        ("dim", C.c_int32), ("precision", C.c_int32), ("integrator", C.c_int32),
        ("eos", C.c_int32), ("aero_case", C.c_int32), ("aero_source", C.c_int32),
        ("use_lam", C.c_int32), ("use_TAB_def", C.c_int32), ("ale", C.c_int32),
        ("max_subits", C.c_int32), ("n_stable_limit", C.c_int32), ("n_unstable_limit", C.c_int32),
    ]

    def copy(self):
        out = Settings()
        C.pointer(out)[0] = self
        return out

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class Derived(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "H", "H_sq", "sr", "W_correc", "W_dx", "B", "dsph_cont", "nu",
        "sim_mass", "bnd_mass", "rho_max", "rho_min", "dx",
        "n_full", "i_n_full", "interp_fac", "A_plate", "A_sphere", "L",
        "sos", "i_sos_sq", "td", "omega", "tmax", "Cdef", "ycoef", "lam_st")]


class Block(C.Structure):
    _fields_ = [
        ("first", C.c_int64), ("last", C.c_int64), ("is_fluid", C.c_int32),
        ("bound_solver", C.c_int32), ("no_slip", C.c_int32), ("n_times", C.c_int32),
        ("vel", C.c_double * 3),
    ]


NO_PLANE = 9999999.0  # default_val: an unset plane


class BlockPlanes(C.Structure):
    _fields_ = [
        ("inlet", C.c_int32), ("fixed_vel_or_dynamic", C.c_int32),
        ("delete_norm", C.c_double * 3), ("delconst", C.c_double),
        ("insert_norm", C.c_double * 3), ("insconst", C.c_double),
        ("aero_norm", C.c_double * 3), ("aeroconst", C.c_double),
    ]


def make_planes(inlet=0, fixed_vel_or_dynamic=0, delete=None, insert=None, aero=None):
    """BlockPlanes with the reference's defaults (unset planes at
    default_val) for the planes not given; a plane is (normal, constant)."""
    q = BlockPlanes()
    q.inlet = int(inlet)
    q.fixed_vel_or_dynamic = int(fixed_vel_or_dynamic)
    for name, plane in (("delete", delete), ("insert", insert), ("aero", aero)):
        nrm, c = plane if plane is not None else ((NO_PLANE,) * 3, NO_PLANE)
        vec = getattr(q, name + "_norm")
        for a in range(3):
            vec[a] = float(nrm[a]) if a < len(nrm) else 0.0
        setattr(q, {"delete": "delconst", "insert": "insconst", "aero": "aeroconst"}[name], float(c))
    return q


class Clock(C.Structure):
    """fjsph_clock: the integrator state carried across steps and restarts."""
    _fields_ = [("current_time", C.c_double), ("last_frame_time", C.c_double), ("delta_t", C.c_double),
                ("cfl", C.c_double), ("current_frame", C.c_int64), ("n_stable", C.c_int64),
                ("n_unstable", C.c_int64), ("next_part_id", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LostInfo(C.Structure):
    """fjsph_lost_info: FindCell's handling of particles in no cell."""
    _fields_ = [("lost", C.c_int64), ("deleted", C.c_int64), ("internal", C.c_int64), ("failed", C.c_int64),
                ("first_cell", C.c_int64), ("first_cell_deleted", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IptSettings(C.Structure):
    """fjsph_ipt_settings (IPT_SETT, src/Var.h:313-331)."""
    _fields_ = [("using_ipt", C.c_int32), ("eq_order", C.c_int32), ("max_x", C.c_double), ("relax", C.c_double),
                ("n_relax", C.c_int32), ("max_steps", C.c_int32)]


def default_ipt(**kw):
    s = IptSettings(using_ipt=0, eq_order=2, max_x=9999999.0, relax=0.6, n_relax=5, max_steps=1000000)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


IPT_SUCCESS, IPT_FAILED, IPT_STOPPED = 1, 2, 3


class IptRecord(C.Structure):
    """fjsph_ipt_record: the end of one IPT track."""
    _fields_ = [("part_id", C.c_int64), ("status", C.c_int32), ("steps", C.c_int32), ("cell", C.c_int32),
                ("face", C.c_int32), ("t0", C.c_double), ("t", C.c_double), ("x0", C.c_double * 3),
                ("x", C.c_double * 3), ("v", C.c_double * 3)]


IPT_DTYPE = np.dtype([("part_id", np.int64), ("status", np.int32), ("steps", np.int32), ("cell", np.int32),
                      ("face", np.int32), ("t0", np.float64), ("t", np.float64), ("x0", np.float64, 3),
                      ("x", np.float64, 3), ("v", np.float64, 3)])


class StepInfo(C.Structure):
    _fields_ = [
        ("rms_error", C.c_double), ("dt", C.c_double), ("cfl", C.c_double),
        ("time", C.c_double), ("max_rho_pc", C.c_double), ("maxf", C.c_double),
        ("maxAf", C.c_double), ("npd", C.c_double), ("iterations", C.c_int32),
        ("rollbacks", C.c_int32),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def default_settings(dim):
    """Reference defaults (src/Var.h:63-310, src/IO.cpp:47-48)."""
    s = Settings()
    s.particle_step = -1.0
    s.H_fac = 2.0
    s.rho_rest = 1000.0
    s.speed_sound = 300.0
    s.gam = 7.0
    s.mu = 8.94e-4
    s.sig = 0.0708
    s.visc_alpha = 0.1
    s.dsph_delta = 0.1
    s.press_back = 0.0
    s.press_pipe = 0.0
    s.rho_var = 50.0
    s.rho_max = 1500.0
    s.rho_min = 500.0
    s.rho_max_iter = 1.0
    g = (0.0, 0.0, -9.81) if dim == 3 else (0.0, -9.81, 0.0)
    for k in range(3):
        s.grav[k] = g[k]
    s.cfl = 1.0
    s.cfl_min = 0.1
    s.cfl_max = 2.0
    s.cfl_step = 0.05
    s.subits_factor = 0.333
    s.min_residual = -7.0
    s.delta_t_max = 1.0
    s.delta_t_min = 0.0
    s.frame_time_interval = 1.0e30
    s.max_shift_vel = 9999999.0
    s.p_ref = 101353.0
    s.rho_g = 1.29251
    s.mu_g = 1.716e-5
    s.temp_g = 298.0
    s.R_g = 287.0
    s.gamma_g = 1.403
    s.lam_cutoff = 0.75
    s.i_interp_fac = 0.5
    s.Cf = 1.0 / 3.0
    s.Ck = 8.0
    s.Cd = 5.0
    s.Cb = 0.5
    s.dim = dim
    s.precision = PREC_FP64
    s.integrator = NEWMARK_BETA
    s.eos = COLE_EOS
    s.aero_case = NO_AERO
    s.aero_source = CONST_VEL
    s.use_lam = 1
    s.use_TAB_def = 0
    s.ale = 0
    s.max_subits = 20
    s.n_stable_limit = 10
    s.n_unstable_limit = 3
    return s


def make_blocks(specs):
    arr = (Block * len(specs))()
    for i, sp in enumerate(specs):
        b = arr[i]
        b.first = int(sp["first"])
        b.last = int(sp["last"])
        b.is_fluid = int(sp.get("is_fluid", 0))
        b.bound_solver = int(sp.get("bound_solver", PRESSURE_G))
        b.no_slip = int(sp.get("no_slip", 0))
        b.n_times = 0
        vel = sp.get("vel", (0.0, 0.0, 0.0))
        for k in range(3):
            b.vel[k] = float(vel[k]) if k < len(vel) else 0.0
    return arr
