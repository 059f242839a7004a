"""Restart files (ctypes over fjsph_amd/libfjsph_restart.so, the calls of
include/fjsph_restart.h).

The file is the reference's `<prefix>_particles.h5` (Write_HDF5 / Read_HDF5,
src/H5IO.cpp:915-1318): root attributes, one group per block ("Boundary k",
"Fluid k") with 1-D datasets of its particles, inlet columns and block
attributes.

    restart.write(engine, "run_particles.h5")          # end of a frame
    eng = Engine(settings)                              # same settings
    restart.read(eng, "run_particles.h5")               # Init_Particles_Restart + Read_HDF5

`write_host` / `read_host` move a host image (a dict of NumPy arrays)
without a GPU.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfjsph_restart.so")
ABI_VERSION = 1
NBUF = 4  # buffer particles per inlet column in the file (src/H5IO.cpp:530-540)

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_LIB = None


class RestartError(RuntimeError):
    pass


class Meta(C.Structure):
    """fjsph_restart_meta: the driver's I/O names and settings."""
    _fields_ = [(name, C.c_char_p) for name in (
        "input_file", "fluid_file", "bound_file", "tau_mesh", "tau_bmap", "tau_sol", "foam_dir", "foam_sol",
        "vlm_file", "output_prefix", "restart_prefix", "output_names", "solver_name", "aero_case_name")] + [
        ("block_names", C.POINTER(C.c_char_p)), ("block_types", _I32), ("particle_order", _I32)] + [
        (name, C.c_double) for name in (
            "scale", "angle_alpha", "hydro_height", "bound_step_factor", "restart_tol", "contangb", "v_ref",
            "M_ref", "max_x_sph", "max_x")] + [("offset_vec", C.c_double * 3)] + [
        (name, C.c_int32) for name in (
            "offset_axis", "single_file", "out_encoding", "max_frames", "init_hydro_pressure", "use_dx",
            "bound_solver", "using_ipt", "ipt_eq_order", "part_out", "streak_out", "cells_out", "foam_is_binary",
            "foam_label_size", "foam_scalar_size", "foam_buoyant_sim", "foam_is_incomp")]


class Header(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n_bound_blocks", C.c_int32), ("n_fluid_blocks", C.c_int32),
                ("bound_points", C.c_int64), ("fluid_points", C.c_int64), ("clock", abi.Clock),
                ("particle_step", C.c_double), ("frame_time_interval", C.c_double), ("scale", C.c_double),
                ("single_file", C.c_int32)]


class Host(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n", C.c_int64), ("n_bound", C.c_int64), ("settings", abi.Settings),
                ("clock", abi.Clock), ("max_points", C.c_int64), ("n_blocks", C.c_int32),
                ("blocks", C.POINTER(abi.Block)), ("planes", C.POINTER(abi.BlockPlanes)),
                ("n_times", _I32), ("times", C.POINTER(_D)), ("vels", C.POINTER(_D)),
                ("n_cols", _I32), ("back", C.POINTER(_I64)), ("buffer", C.POINTER(_I64)),
                ("x", _D), ("v", _D), ("acc", _D), ("p", _D), ("rho", _D), ("rrho", _D), ("m", _D),
                ("b", _I32), ("pid", _I64), ("cell_id", _I64), ("cell_v", _D), ("cell_p", _D), ("cell_rho", _D)]


SIGNATURES = {
    "fjsph_restart_abi_version": (C.c_int32, []),
    "fjsph_restart_last_error": (C.c_char_p, []),
    "fjsph_restart_meta_default": (None, [C.POINTER(Meta), C.c_int32]),
    "fjsph_restart_read_header": (C.c_int, [C.c_char_p, C.POINTER(Header)]),
    "fjsph_restart_write_host": (C.c_int, [C.c_char_p, C.POINTER(Host), C.POINTER(Meta)]),
    "fjsph_restart_read_host": (C.c_int, [C.c_char_p, C.POINTER(Host)]),
    "fjsph_restart_write": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(Meta)]),
    "fjsph_restart_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(abi.Block),
                                     C.POINTER(abi.BlockPlanes)]),
}

FIELDS_D = ("x", "v", "acc", "cell_v")      # dim per particle
FIELDS_1 = ("p", "rho", "rrho", "m", "cell_p", "cell_rho")


def load():
    """Load the restart library; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RestartError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    from . import engine
    engine.load()  # the engine library first, from this package
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fjsph_restart_abi_version() != ABI_VERSION:
        raise RestartError("restart library ABI version mismatch")
    _LIB = lib
    return lib


def _chk(rc, what):
    if rc != abi.OK:
        raise RestartError(f"{what}: status {rc}: {load().fjsph_restart_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(t)


def default_meta(dim, **kw):
    """fjsph_restart_meta with the reference's defaults; keyword strings and
    numbers override fields."""
    m = Meta()
    load().fjsph_restart_meta_default(C.byref(m), int(dim))
    keep = []
    for k, v in kw.items():
        if isinstance(v, str):
            b = v.encode()
            keep.append(b)
            setattr(m, k, b)
        else:
            setattr(m, k, v)
    m._keep = keep
    return m


def read_header(path):
    h = Header()
    _chk(load().fjsph_restart_read_header(os.fsencode(path), C.byref(h)), "read_header")
    return h


def write(engine, path, meta=None):
    """Write the engine's state (pnp1, blocks, planes, motion, inlets,
    clock) as a restart file."""
    _chk(load().fjsph_restart_write(engine.h, os.fsencode(path), C.byref(meta) if meta is not None else None),
         "restart_write")


def read(engine, path, blocks=None, planes=None):
    """Load a restart file into an engine created with the run's settings.
    blocks / planes: the block definitions (lists of abi.Block /
    abi.BlockPlanes), or None to take them from the file."""
    nb = 0
    barr = parr = None
    if blocks is not None:
        nb = len(blocks)
        barr = (abi.Block * nb)(*blocks)
    if planes is not None:
        parr = (abi.BlockPlanes * len(planes))(*planes)
    _chk(load().fjsph_restart_read(engine.h, os.fsencode(path), nb, barr, parr), "restart_read")


def _host_struct(img, keep):
    """Host struct pointing at the arrays of `img` (kept alive in `keep`)."""
    st = Host()
    dim = int(img["dim"])
    n = int(img["n"])
    st.dim, st.n, st.n_bound = dim, n, int(img["n_bound"])
    st.settings = img.get("settings", abi.default_settings(dim))
    st.clock = img["clock"]
    st.max_points = int(img.get("max_points", 9999999))
    blocks, planes = img["blocks"], img["planes"]
    nbk = len(blocks)
    if len(planes) != nbk:
        raise RestartError("one planes entry per block")
    st.n_blocks = nbk
    ba = (abi.Block * nbk)(*blocks)
    pa = (abi.BlockPlanes * nbk)(*planes)
    keep += [ba, pa]
    st.blocks = C.cast(ba, C.POINTER(abi.Block))
    st.planes = C.cast(pa, C.POINTER(abi.BlockPlanes))
    nt = np.array([len(t) for t in img.get("times", [[]] * nbk)], np.int32)
    nc = np.array([0 if b is None else len(b) for b in img.get("back", [None] * nbk)], np.int32)
    keep += [nt, nc]
    st.n_times, st.n_cols = _p(nt, _I32), _p(nc, _I32)
    tp, vp, bp, fp = (_D * nbk)(), (_D * nbk)(), (_I64 * nbk)(), (_I64 * nbk)()
    keep += [tp, vp, bp, fp]
    for k in range(nbk):
        if nt[k]:
            t = np.ascontiguousarray(img["times"][k], np.float64)
            v = np.ascontiguousarray(img["vels"][k], np.float64)
            if v.size != 3 * t.size:
                raise RestartError(f"block {k}: vels needs 3 values per time")
            keep += [t, v]
            tp[k], vp[k] = _p(t, _D), _p(v, _D)
        if nc[k]:
            bk = np.ascontiguousarray(img["back"][k], np.int64)
            bf = np.ascontiguousarray(img["buffer"][k], np.int64)
            if bf.size != NBUF * bk.size:
                raise RestartError(f"block {k}: the restart layout holds {NBUF} buffer particles per inlet column")
            keep += [bk, bf]
            bp[k], fp[k] = _p(bk, _I64), _p(bf, _I64)
    st.times, st.vels, st.back, st.buffer = tp, vp, bp, fp
    for f in FIELDS_D + FIELDS_1:
        a = np.ascontiguousarray(img[f], np.float64)
        if a.size != n * (dim if f in FIELDS_D else 1):
            raise RestartError(f"{f}: {a.size} values for {n} particles")
        keep.append(a)
        setattr(st, f, _p(a, _D))
    b = np.ascontiguousarray(img["b"], np.int32)
    pid = np.ascontiguousarray(img["pid"], np.int64)
    cid = np.ascontiguousarray(img["cell_id"], np.int64)
    for name, a in (("b", b), ("pid", pid), ("cell_id", cid)):
        if a.size != n:
            raise RestartError(f"{name}: {a.size} values for {n} particles")
    keep += [b, pid, cid]
    st.b, st.pid, st.cell_id = _p(b, _I32), _p(pid, _I64), _p(cid, _I64)
    return st


def write_host(path, img, meta=None):
    """Write a host image: dict with dim, n, n_bound, clock (abi.Clock),
    settings (abi.Settings), max_points, blocks, planes, optional times /
    vels / back / buffer per block, and the particle arrays (x, v, acc,
    cell_v: (n, dim); p, rho, rrho, m, cell_p, cell_rho, b, pid, cell_id)."""
    keep = []
    st = _host_struct(img, keep)
    _chk(load().fjsph_restart_write_host(os.fsencode(path), C.byref(st), C.byref(meta) if meta is not None else None),
         "restart_write_host")


def read_host(path):
    """The host image of a restart file (the dict write_host takes, with
    settings.particle_step / frame_time_interval from the file)."""
    h = read_header(path)
    nbk = h.n_bound_blocks + h.n_fluid_blocks
    st = Host()
    ba = (abi.Block * nbk)()
    pa = (abi.BlockPlanes * nbk)()
    nt = np.zeros(nbk, np.int32)
    nc = np.zeros(nbk, np.int32)
    st.n_blocks = nbk
    st.blocks = C.cast(ba, C.POINTER(abi.Block))
    st.planes = C.cast(pa, C.POINTER(abi.BlockPlanes))
    st.n_times, st.n_cols = _p(nt, _I32), _p(nc, _I32)
    _chk(load().fjsph_restart_read_host(os.fsencode(path), C.byref(st)), "restart_read_host")  # sizes
    dim, n = int(st.dim), int(st.n)
    img = {"dim": dim, "n": n, "n_bound": int(st.n_bound), "clock": st.clock, "settings": st.settings,
           "max_points": int(st.max_points), "blocks": list(ba), "planes": list(pa)}
    for f in FIELDS_D:
        img[f] = np.zeros((n, dim))
    for f in FIELDS_1:
        img[f] = np.zeros(n)
    img["b"] = np.zeros(n, np.int32)
    img["pid"] = np.zeros(n, np.int64)
    img["cell_id"] = np.zeros(n, np.int64)
    img["times"] = [np.zeros(int(nt[k])) for k in range(nbk)]
    img["vels"] = [np.zeros(3 * int(nt[k])) for k in range(nbk)]
    img["back"] = [np.zeros(int(nc[k]), np.int64) if nc[k] else None for k in range(nbk)]
    img["buffer"] = [np.zeros(NBUF * int(nc[k]), np.int64) if nc[k] else None for k in range(nbk)]
    keep = []
    st2 = _host_struct(img, keep)
    st2.n_blocks = nbk
    _chk(load().fjsph_restart_read_host(os.fsencode(path), C.byref(st2)), "restart_read_host")
    img["clock"] = st2.clock
    img["settings"] = st2.settings
    img["blocks"] = list(C.cast(st2.blocks, C.POINTER(abi.Block * nbk)).contents)
    img["planes"] = list(C.cast(st2.planes, C.POINTER(abi.BlockPlanes * nbk)).contents)
    img["max_points"] = int(st2.max_points)
    return img
