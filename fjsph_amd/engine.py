"""Python host side of the WCSPH step engine: a ctypes binding of the C-ABI
in include/fjsph_engine.h, exposing the reference's per-step units
(update_neighbours, dSPH_PreStep, Detect_Surface, dissipation_terms,
particle_shift, get_acc_and_Rrho, Do_NB_Iter, find_timestep,
Integrator::integrate) on a HIP device.

There is no CPU fallback: if libfjsph_engine.so is missing or no HIP device
is visible, construction raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfjsph_engine.so")
_LIB = None

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_F32 = C.POINTER(C.c_float)
_VP = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/fjsph_engine.h
SIGNATURES = {
    "fjsph_abi_version": (C.c_int32, []),
    "fjsph_settings_default": (None, [C.POINTER(abi.Settings), C.c_int32]),
    "fjsph_derive": (C.c_int, [C.POINTER(abi.Settings), C.POINTER(abi.Derived)]),
    "fjsph_settings_parse_file": (C.c_int, [C.POINTER(abi.Settings), C.c_char_p]),
    "fjsph_create": (C.c_int, [C.POINTER(abi.Settings), C.c_int32, C.POINTER(_VP)]),
    "fjsph_destroy": (None, [_VP]),
    "fjsph_last_error": (C.c_char_p, [_VP]),
    "fjsph_upload_state": (C.c_int, [_VP, C.c_int64, C.c_int64, _D, _D, _D, _D, _D, _I32]),
    "fjsph_set_blocks": (C.c_int, [_VP, C.c_int32, C.POINTER(abi.Block)]),
    "fjsph_set_block_motion": (C.c_int, [_VP, C.c_int32, C.c_int32, _D, _D]),
    "fjsph_download_field": (C.c_int, [_VP, C.c_int32, C.c_int32, _D]),
    "fjsph_upload_field": (C.c_int, [_VP, C.c_int32, C.c_int32, _D]),
    "fjsph_build_neighbours": (C.c_int, [_VP]),
    "fjsph_get_neighbour_counts": (C.c_int, [_VP, _I32]),
    "fjsph_get_neighbours": (C.c_int, [_VP, _I64, _I64, C.c_int64]),
    "fjsph_dsph_prestep": (C.c_int, [_VP, _D]),
    "fjsph_aero_velocity": (C.c_int, [_VP]),
    "fjsph_detect_surface": (C.c_int, [_VP]),
    "fjsph_dissipation_terms": (C.c_int, [_VP]),
    "fjsph_particle_shift": (C.c_int, [_VP]),
    "fjsph_boundary_update": (C.c_int, [_VP]),
    "fjsph_acc_rrho": (C.c_int, [_VP, C.c_double]),
    "fjsph_nb_iter": (C.c_int, [_VP, C.c_double, _D]),
    "fjsph_find_timestep": (C.c_int, [_VP, _D]),
    "fjsph_integrate_no_update": (C.c_int, [_VP, C.POINTER(abi.StepInfo)]),
    "fjsph_integrate_step": (C.c_int, [_VP, C.POINTER(abi.StepInfo)]),
    "fjsph_advance_frame": (C.c_int, [_VP]),
    "fjsph_get_time": (C.c_int, [_VP, _D, _D, _D]),
    "fjsph_set_time": (C.c_int, [_VP, C.c_double, C.c_double, C.c_double]),
    "fjsph_enable_timers": (C.c_int, [_VP, C.c_int32]),
    "fjsph_get_timers": (C.c_int, [_VP, _D, _I64]),
    "fjsph_pass_bytes": (C.c_int, [_VP, C.c_int32, _D]),
    "fjsph_neighbour_stats": (C.c_int, [_VP, _I64, _I32]),
    "fjsph_set_mesh": (C.c_int, [_VP, C.c_int64, _D, C.c_int64, _I64, _I64, _I64, C.c_int64, _I64, _I64,
                                 _D, _D, _D]),
    "fjsph_get_mesh_cells": (C.c_int, [_VP, _D, _D, _D, _D]),
    "fjsph_mesh_lost": (C.c_int, [_VP, _I64]),
    "fjsph_mesh_lost_info": (C.c_int, [_VP, C.POINTER(abi.LostInfo)]),
    "fjsph_mesh_eval": (C.c_int, [_VP, C.c_int64, _D, _I32, _D]),
    "fjsph_set_surrogate": (C.c_int, [_VP, C.c_int32, _F32, _F32, _D, _D, _D, _D]),
    "fjsph_surrogate_eval": (C.c_int, [_VP, C.c_int64, _D, _D]),
    "fjsph_surrogate_train": (C.c_int, [_VP, C.c_int64, _D, _D, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                        C.c_uint64, _D, _D, _D, _D, _F32, _F32, _D]),
    "fjsph_surrogate_grad": (C.c_int, [_VP, C.c_int32, C.c_int32, _F32, _F32, _F32, _F32, _F32, _F32, _D]),
    "fjsph_set_groups": (C.c_int, [_VP, C.c_int32, _I32]),
    "fjsph_get_group_info": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_void_p]),
    # device pointers (no allocation inside)
    "fjsph_surrogate_grad_noalloc": (C.c_int, [_VP, C.c_int32, C.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "fjsph_set_block_planes": (C.c_int, [_VP, C.c_int32, C.POINTER(abi.BlockPlanes)]),
    "fjsph_set_inlet": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _I64, _I64]),
    "fjsph_set_max_points": (C.c_int, [_VP, C.c_int64]),
    "fjsph_get_counts": (C.c_int, [_VP, _I64, _I64]),
    "fjsph_get_block_range": (C.c_int, [_VP, C.c_int32, _I64, _I64]),
    "fjsph_get_blocks": (C.c_int, [_VP, _I32, C.POINTER(abi.Block), C.c_int32]),
    "fjsph_get_block_motion": (C.c_int, [_VP, C.c_int32, _I32, _D, _D]),
    "fjsph_get_block_planes": (C.c_int, [_VP, C.c_int32, C.POINTER(abi.BlockPlanes)]),
    "fjsph_get_inlet": (C.c_int, [_VP, C.c_int32, _I32, _I32, _I64, _I64]),
    "fjsph_get_clock": (C.c_int, [_VP, C.POINTER(abi.Clock)]),
    "fjsph_set_clock": (C.c_int, [_VP, C.POINTER(abi.Clock)]),
    "fjsph_get_settings": (C.c_int, [_VP, C.POINTER(abi.Settings)]),
    "fjsph_get_max_points": (C.c_int, [_VP, _I64]),
    "fjsph_ipt_settings_default": (None, [C.POINTER(abi.IptSettings)]),
    "fjsph_set_ipt": (C.c_int, [_VP, C.POINTER(abi.IptSettings)]),
    "fjsph_ipt_track": (C.c_int, [_VP, C.c_int64, _D, _D, _I32, _I64, C.c_double, C.c_void_p]),
    "fjsph_ipt_count": (C.c_int, [_VP, _I64, _I64, _I64]),
    "fjsph_ipt_records": (C.c_int, [_VP, C.c_int64, C.c_int64, C.c_void_p]),
    "fjsph_ipt_clear": (C.c_int, [_VP]),
    "fjsph_set_block_targets": (C.c_int, [_VP, C.c_int32]),
    "fjsph_set_reorder_interval": (C.c_int, [_VP, C.c_int32]),
    "fjsph_get_block_targets": (C.c_int, [_VP, _I32]),
    "fjsph_autotune_blocks": (C.c_int, [_VP, C.c_int32, _I32, C.c_int32, _D, _I32]),
}


class EngineError(RuntimeError):
    pass


def load():
    """Load libfjsph_engine.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise EngineError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.fjsph_abi_version() != abi.ABI_VERSION:
        raise EngineError("engine ABI version mismatch")
    _LIB = lib
    return lib


def _ptr(a, t=_D):
    return a.ctypes.data_as(t)


def settings_default(dim):
    s = abi.Settings()
    load().fjsph_settings_default(C.byref(s), dim)
    return s


def derive(settings):
    d = abi.Derived()
    rc = load().fjsph_derive(C.byref(settings), C.byref(d))
    if rc != abi.OK:
        raise EngineError(f"fjsph_derive failed ({rc})")
    return d


def parse_para(path, dim):
    s = settings_default(dim)
    rc = load().fjsph_settings_parse_file(C.byref(s), path.encode())
    if rc != abi.OK:
        raise EngineError(f"could not parse {path} ({rc})")
    return s


class Engine:
    """One engine context on a HIP device; methods follow the reference units."""

    def __init__(self, case_or_settings, device=0):
        self.lib = load()
        settings = case_or_settings.settings if hasattr(case_or_settings, "settings") else case_or_settings
        self.settings = settings.copy()
        self.dim = int(settings.dim)
        h = _VP()
        rc = self.lib.fjsph_create(C.byref(self.settings), int(device), C.byref(h))
        if rc != abi.OK:
            raise EngineError(f"fjsph_create failed with status {rc}")
        self.h = h
        if hasattr(case_or_settings, "x"):
            self.upload_case(case_or_settings)

    @property
    def n(self):
        """Current particle count (inlets add particles, deletion removes them)."""
        return self.counts()[0]

    def counts(self):
        nt, nb = np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.fjsph_get_counts(self.h, _ptr(nt, _I64), _ptr(nb, _I64)), "get_counts")
        return int(nt[0]), int(nb[0])

    def block_range(self, block):
        f, l = np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.fjsph_get_block_range(self.h, int(block), _ptr(f, _I64), _ptr(l, _I64)), "block_range")
        return int(f[0]), int(l[0])

    def set_block_planes(self, block, planes):
        """planes: an abi.BlockPlanes (abi.make_planes)."""
        self._planes = planes
        self._chk(self.lib.fjsph_set_block_planes(self.h, int(block), C.byref(planes)), "set_block_planes")

    def set_inlet(self, block, back, buffer):
        """back: host indices of the columns' BACK particles (n_cols);
        buffer: (n_cols, n_buf) host indices of their BUFFER particles."""
        bk = np.ascontiguousarray(back, np.int64)
        bf = np.ascontiguousarray(buffer, np.int64).reshape(bk.size, -1)
        self._chk(self.lib.fjsph_set_inlet(self.h, int(block), bk.size, bf.shape[1], _ptr(bk, _I64), _ptr(bf, _I64)),
                  "set_inlet")

    def inlet(self, block):
        """(back, buffer) of the block's inlet columns in host indices, or
        None when the block has no inlet."""
        nc, nb = np.zeros(1, np.int32), np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_get_inlet(self.h, int(block), _ptr(nc, _I32), _ptr(nb, _I32), None, None), "get_inlet")
        if nc[0] == 0:
            return None
        bk = np.zeros(int(nc[0]), np.int64)
        bf = np.zeros((int(nc[0]), int(nb[0])), np.int64)
        self._chk(self.lib.fjsph_get_inlet(self.h, int(block), _ptr(nc, _I32), _ptr(nb, _I32), _ptr(bk, _I64),
                                           _ptr(bf, _I64)), "get_inlet")
        return bk, bf

    def blocks(self):
        nb = np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_get_blocks(self.h, _ptr(nb, _I32), None, 0), "get_blocks")
        arr = (abi.Block * int(nb[0]))()
        self._chk(self.lib.fjsph_get_blocks(self.h, _ptr(nb, _I32), arr, int(nb[0])), "get_blocks")
        return list(arr)

    def block_planes(self, block):
        pl = abi.BlockPlanes()
        self._chk(self.lib.fjsph_get_block_planes(self.h, int(block), C.byref(pl)), "get_block_planes")
        return pl

    def clock(self):
        c = abi.Clock()
        self._chk(self.lib.fjsph_get_clock(self.h, C.byref(c)), "get_clock")
        return c

    def set_clock(self, clock):
        self._chk(self.lib.fjsph_set_clock(self.h, C.byref(clock)), "set_clock")

    def settings(self):
        st = abi.Settings()
        self._chk(self.lib.fjsph_get_settings(self.h, C.byref(st)), "get_settings")
        return st

    def set_block_targets(self, targets):
        self._chk(self.lib.fjsph_set_block_targets(self.h, int(targets)), "set_block_targets")

    def get_block_targets(self):
        v = np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_get_block_targets(self.h, _ptr(v, _I32)), "get_block_targets")
        return int(v[0])

    def set_reorder_interval(self, steps):
        """fjsph_set_reorder_interval: cell-order permutation every `steps` steps."""
        self._chk(self.lib.fjsph_set_reorder_interval(self.h, int(steps)), "set_reorder_interval")

    def block_targets(self):
        t = np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_get_block_targets(self.h, _ptr(t, _I32)), "get_block_targets")
        return int(t[0])

    def autotune_blocks(self, candidates=(128, 192, 256, 384, 512), reps=3):
        """fjsph_autotune_blocks: (best targets per block, {candidate: ms})."""
        c = np.ascontiguousarray(candidates, np.int32)
        ms = np.zeros(c.size)
        best = np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_autotune_blocks(self.h, c.size, _ptr(c, _I32), int(reps), _ptr(ms), _ptr(best, _I32)),
                  "autotune_blocks")
        return int(best[0]), dict(zip(c.tolist(), ms.tolist()))

    def set_ipt(self, settings):
        """settings: abi.IptSettings (abi.default_ipt)."""
        self._chk(self.lib.fjsph_set_ipt(self.h, C.byref(settings)), "set_ipt")

    def ipt_track(self, x, v, cell, part_id=None, t0=0.0):
        """Track particles through the mesh; a structured array of records
        (abi.IPT_DTYPE)."""
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        c = np.ascontiguousarray(cell, np.int32)
        n = c.size
        pid = None if part_id is None else np.ascontiguousarray(part_id, np.int64)
        out = np.zeros(n, abi.IPT_DTYPE)
        self._chk(self.lib.fjsph_ipt_track(self.h, n, _ptr(x), _ptr(v), _ptr(c, _I32),
                                           None if pid is None else _ptr(pid, _I64), float(t0),
                                           out.ctypes.data_as(C.c_void_p)), "ipt_track")
        return out

    def ipt_count(self):
        a, b, c = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.fjsph_ipt_count(self.h, _ptr(a, _I64), _ptr(b, _I64), _ptr(c, _I64)), "ipt_count")
        return int(a[0]), int(b[0]), int(c[0])

    def ipt_records(self):
        n = self.ipt_count()[0]
        out = np.zeros(n, abi.IPT_DTYPE)
        self._chk(self.lib.fjsph_ipt_records(self.h, 0, n, out.ctypes.data_as(C.c_void_p)), "ipt_records")
        return out

    def ipt_clear(self):
        self._chk(self.lib.fjsph_ipt_clear(self.h), "ipt_clear")

    def set_max_points(self, max_points):
        self._chk(self.lib.fjsph_set_max_points(self.h, int(max_points)), "set_max_points")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fjsph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc, what):
        if rc != abi.OK:
            msg = self.lib.fjsph_last_error(self.h)
            raise EngineError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    # ---- state
    def upload_case(self, case):
        self.set_blocks(case.blocks)
        self.upload_state(case.x, case.v, case.rho, case.p, case.m, case.b, case.n_bound)
        if getattr(case, "mesh", None) is not None:
            self.set_mesh(case.mesh)
        for blk, kw in (getattr(case, "planes", None) or {}).items():
            self.set_block_planes(blk, abi.make_planes(**kw))
        for blk, back, buffer in getattr(case, "inlets", None) or []:
            self.set_inlet(blk, back, buffer)
        if getattr(case, "max_points", None):
            self.set_max_points(case.max_points)

    def upload_state(self, x, v, rho, p, m, b, n_bound):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        m = np.ascontiguousarray(m, np.float64)
        b = np.ascontiguousarray(b, np.int32)
        self._chk(self.lib.fjsph_upload_state(self.h, x.shape[0], int(n_bound), _ptr(x), _ptr(v), _ptr(rho),
                                              _ptr(p), _ptr(m), _ptr(b, _I32)), "upload_state")

    def set_blocks(self, specs):
        arr = abi.make_blocks(specs)
        self._chk(self.lib.fjsph_set_blocks(self.h, len(specs), arr), "set_blocks")

    def set_block_motion(self, block, times, vels):
        t = np.ascontiguousarray(times, np.float64)
        v = np.ascontiguousarray(vels, np.float64).reshape(-1, 3)
        self._chk(self.lib.fjsph_set_block_motion(self.h, int(block), len(t), _ptr(t), _ptr(v)), "set_block_motion")

    def field(self, name, which=abi.STATE_NP1):
        w = abi.field_width(name, self.dim)
        n = self.n
        out = np.zeros(n * w)
        self._chk(self.lib.fjsph_download_field(self.h, which, abi.F[name], _ptr(out)), f"download {name}")
        return out.reshape(n, w) if w > 1 else out

    def set_field(self, name, values, which=abi.STATE_NP1):
        a = np.ascontiguousarray(values, np.float64).ravel()
        self._chk(self.lib.fjsph_upload_field(self.h, which, abi.F[name], _ptr(a)), f"upload {name}")

    # ---- passes
    def build_neighbours(self):
        self._chk(self.lib.fjsph_build_neighbours(self.h), "build_neighbours")

    def neighbour_counts(self):
        out = np.zeros(self.n, np.int32)
        self._chk(self.lib.fjsph_get_neighbour_counts(self.h, _ptr(out, _I32)), "neighbour_counts")
        return out

    def neighbours(self):
        counts = self.neighbour_counts()
        total = int(counts.astype(np.int64).sum())
        off = np.zeros(self.n + 1, np.int64)
        idx = np.zeros(total, np.int64)
        self._chk(self.lib.fjsph_get_neighbours(self.h, _ptr(off, _I64), _ptr(idx, _I64), total), "neighbours")
        return off, idx

    def set_mesh(self, mesh):
        """fjsph_set_mesh with a cases.Mesh."""
        self._mesh = mesh
        a = lambda x, t=np.float64: np.ascontiguousarray(x, dtype=t)
        self._mesh_arrays = [a(mesh.verts), a(mesh.face_off, np.int64), a(mesh.face_verts, np.int64),
                             a(mesh.face_lr, np.int64), a(mesh.cell_off, np.int64), a(mesh.cell_faces, np.int64),
                             a(mesh.vert_vel), a(mesh.vert_p), a(mesh.vert_rho)]
        v, fo, fv, lr, co, cf, vv, vp, vr = self._mesh_arrays
        self._chk(self.lib.fjsph_set_mesh(self.h, mesh.nverts, _ptr(v), mesh.nfaces, _ptr(fo, _I64), _ptr(fv, _I64),
                                          _ptr(lr, _I64), mesh.ncells, _ptr(co, _I64), _ptr(cf, _I64), _ptr(vv),
                                          _ptr(vp), _ptr(vr)), "set_mesh")

    def mesh_cells(self):
        nc = self._mesh.ncells
        c, v = np.zeros((nc, self.dim)), np.zeros((nc, self.dim))
        p, r = np.zeros(nc), np.zeros(nc)
        self._chk(self.lib.fjsph_get_mesh_cells(self.h, _ptr(c), _ptr(v), _ptr(p), _ptr(r)), "get_mesh_cells")
        return c, v, p, r

    def mesh_eval(self, points):
        """The mesh field at points: (cells, vals)."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        cells = np.zeros(pts.shape[0], np.int32)
        vals = np.zeros((pts.shape[0], self.dim + 2))
        self._chk(self.lib.fjsph_mesh_eval(self.h, pts.shape[0], _ptr(pts), _ptr(cells, _I32), _ptr(vals)),
                  "mesh_eval")
        return cells, vals

    def mesh_lost(self):
        v = np.zeros(1, np.int64)
        self._chk(self.lib.fjsph_mesh_lost(self.h, _ptr(v, _I64)), "mesh_lost")
        return int(v[0])

    def lost_info(self):
        """fjsph_mesh_lost_info as a dict (lost, deleted, internal, failed)."""
        r = abi.LostInfo()
        self._chk(self.lib.fjsph_mesh_lost_info(self.h, C.byref(r)), "mesh_lost_info")
        return r.as_dict()

    def set_surrogate(self, sg):
        """fjsph_set_surrogate with a surrogate.Surrogate (or dict of the same fields)."""
        g = (lambda k: getattr(sg, k)) if not isinstance(sg, dict) else sg.__getitem__
        self._sg_arrays = [np.ascontiguousarray(g("weights"), dtype=np.float32),
                           np.ascontiguousarray(g("biases"), dtype=np.float32),
                           np.ascontiguousarray(g("in_shift"), dtype=np.float64),
                           np.ascontiguousarray(g("in_scale"), dtype=np.float64),
                           np.ascontiguousarray(g("out_shift"), dtype=np.float64),
                           np.ascontiguousarray(g("out_scale"), dtype=np.float64)]
        w, b, ish, isc, osh, osc = self._sg_arrays
        self._chk(self.lib.fjsph_set_surrogate(self.h, int(g("n_hidden")), _ptr(w, _F32), _ptr(b, _F32), _ptr(ish),
                                               _ptr(isc), _ptr(osh), _ptr(osc)), "set_surrogate")

    def surrogate_eval(self, points):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        out = np.zeros((pts.shape[0], self.dim + 2))
        self._chk(self.lib.fjsph_surrogate_eval(self.h, pts.shape[0], _ptr(pts), _ptr(out)), "surrogate_eval")
        return out

    def surrogate_train(self, points, values, n_hidden, weights, biases, in_shift, in_scale, out_shift, out_scale,
                        steps, batch, lr, seed):
        """fjsph_surrogate_train: returns (weights, biases, per-step loss)."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        vals = np.ascontiguousarray(values, dtype=np.float64)
        w = np.array(weights, dtype=np.float32)
        b = np.array(biases, dtype=np.float32)
        sh = [np.ascontiguousarray(a, dtype=np.float64) for a in (in_shift, in_scale, out_shift, out_scale)]
        loss = np.zeros(max(int(steps), 1))
        self._chk(self.lib.fjsph_surrogate_train(self.h, pts.shape[0], _ptr(pts), _ptr(vals), int(n_hidden),
                                                 int(steps), int(batch), float(lr), int(seed) & (2 ** 64 - 1),
                                                 _ptr(sh[0]), _ptr(sh[1]), _ptr(sh[2]), _ptr(sh[3]), _ptr(w, _F32),
                                                 _ptr(b, _F32), _ptr(loss)), "surrogate_train")
        return w, b, loss[:int(steps)]

    def surrogate_grad(self, n_hidden, weights, biases, xn, tn):
        """fjsph_surrogate_grad: (dweights, dbiases, loss) of one batch of
        normalised samples (batch a multiple of 64)."""
        x = np.ascontiguousarray(xn, dtype=np.float32)
        t = np.ascontiguousarray(tn, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        b = np.ascontiguousarray(biases, dtype=np.float32)
        dw, db = np.zeros_like(w), np.zeros_like(b)
        loss = np.zeros(1)
        self._chk(self.lib.fjsph_surrogate_grad(self.h, int(n_hidden), x.shape[0], _ptr(x, _F32), _ptr(t, _F32),
                                                _ptr(w, _F32), _ptr(b, _F32), _ptr(dw, _F32), _ptr(db, _F32),
                                                _ptr(loss)), "surrogate_grad")
        return dw, db, float(loss[0])

    def set_groups(self, group):
        """fjsph_set_groups: independent simulations, group[i] of host
        particle i (0 .. n_groups - 1)."""
        g = np.ascontiguousarray(group, dtype=np.int32)
        ng = int(g.max()) + 1 if g.size else 0
        self._chk(self.lib.fjsph_set_groups(self.h, ng, _ptr(g, _I32)), "set_groups")
        self.n_groups = ng

    def group_info(self):
        """Per-group step information of the last step (list of dicts)."""
        ng = getattr(self, "n_groups", 0)
        arr = (abi.StepInfo * max(ng, 1))()
        self._chk(self.lib.fjsph_get_group_info(self.h, 0, ng, C.cast(arr, C.c_void_p)), "get_group_info")
        return [arr[k].as_dict() for k in range(ng)]

    def neighbour_stats(self):
        pairs = np.zeros(1, np.int64)
        mx = np.zeros(1, np.int32)
        self._chk(self.lib.fjsph_neighbour_stats(self.h, _ptr(pairs, _I64), _ptr(mx, _I32)), "neighbour_stats")
        return int(pairs[0]), int(mx[0])

    def dsph_prestep(self):
        v = np.zeros(1)
        self._chk(self.lib.fjsph_dsph_prestep(self.h, _ptr(v)), "dsph_prestep")
        return float(v[0])

    def aero_velocity(self):
        self._chk(self.lib.fjsph_aero_velocity(self.h), "aero_velocity")

    def detect_surface(self):
        self._chk(self.lib.fjsph_detect_surface(self.h), "detect_surface")

    def dissipation_terms(self):
        self._chk(self.lib.fjsph_dissipation_terms(self.h), "dissipation_terms")

    def particle_shift(self):
        self._chk(self.lib.fjsph_particle_shift(self.h), "particle_shift")

    def boundary_update(self):
        self._chk(self.lib.fjsph_boundary_update(self.h), "boundary_update")

    def acc_rrho(self, npd):
        self._chk(self.lib.fjsph_acc_rrho(self.h, float(npd)), "acc_rrho")

    def nb_iter(self, npd):
        v = np.zeros(1)
        self._chk(self.lib.fjsph_nb_iter(self.h, float(npd), _ptr(v)), "nb_iter")
        return float(v[0])

    def find_timestep(self):
        v = np.zeros(1)
        self._chk(self.lib.fjsph_find_timestep(self.h, _ptr(v)), "find_timestep")
        return float(v[0])

    def integrate_no_update(self):
        info = abi.StepInfo()
        self._chk(self.lib.fjsph_integrate_no_update(self.h, C.byref(info)), "integrate_no_update")
        return info.as_dict()

    def integrate_step(self):
        info = abi.StepInfo()
        self._chk(self.lib.fjsph_integrate_step(self.h, C.byref(info)), "integrate_step")
        return info.as_dict()

    def advance_frame(self):
        self._chk(self.lib.fjsph_advance_frame(self.h), "advance_frame")

    def get_time(self):
        t, dt, cfl = np.zeros(1), np.zeros(1), np.zeros(1)
        self._chk(self.lib.fjsph_get_time(self.h, _ptr(t), _ptr(dt), _ptr(cfl)), "get_time")
        return float(t[0]), float(dt[0]), float(cfl[0])

    def set_time(self, t, dt, cfl):
        self._chk(self.lib.fjsph_set_time(self.h, float(t), float(dt), float(cfl)), "set_time")

    # ---- measurement
    def enable_timers(self, on=True):
        self._chk(self.lib.fjsph_enable_timers(self.h, int(bool(on))), "enable_timers")

    def timers(self):
        ms = np.zeros(len(abi.TIMER_NAMES))
        cnt = np.zeros(len(abi.TIMER_NAMES), np.int64)
        self._chk(self.lib.fjsph_get_timers(self.h, _ptr(ms), _ptr(cnt, _I64)), "get_timers")
        return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(abi.TIMER_NAMES)}

    def pass_bytes(self, pass_id):
        v = np.zeros(1)
        self._chk(self.lib.fjsph_pass_bytes(self.h, int(pass_id), _ptr(v)), "pass_bytes")
        return float(v[0])
