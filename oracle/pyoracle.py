"""ctypes front end of the FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It drives liboracle.so (or libbaseline.so for timing) with the
same arrays and settings the engine takes.
"""
import ctypes as C
import os

import numpy as np

from fjsph_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


def _ptr(a, t=_D):
    return a.ctypes.data_as(t)


def load(kind="oracle"):
    """kind 'oracle' (the checker build) or 'baseline' (reference flags)."""
    if kind in _LIBS:
        return _LIBS[kind]
    path = os.path.join(_HERE, "liboracle.so" if kind == "oracle" else "libbaseline.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() or make -C oracle")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "orc_create": (vp, [C.POINTER(abi.Settings), C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_derive": (C.c_int, [C.POINTER(abi.Settings), C.POINTER(abi.Derived)]),
        "orc_set_serial_dissipation": (C.c_int, [vp, C.c_int]),
        "orc_upload_state": (C.c_int, [vp, C.c_int64, C.c_int64, _D, _D, _D, _D, _D, _I32]),
        "orc_set_blocks": (C.c_int, [vp, C.c_int, C.POINTER(abi.Block)]),
        "orc_set_block_motion": (C.c_int, [vp, C.c_int, C.c_int, _D, _D]),
        "orc_set_block_planes": (C.c_int, [vp, C.c_int, C.POINTER(abi.BlockPlanes)]),
        "orc_set_inlet": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _I64, _I64]),
        "orc_set_max_points": (C.c_int, [vp, C.c_int64]),
        "orc_get_counts": (C.c_int, [vp, _I64, _I64]),
        "orc_get_block_range": (C.c_int, [vp, C.c_int, _I64, _I64]),
        "orc_download_field": (C.c_int, [vp, C.c_int, C.c_int, _D]),
        "orc_upload_field": (C.c_int, [vp, C.c_int, C.c_int, _D]),
        "orc_build_neighbours": (C.c_int, [vp, C.c_int]),
        "orc_neighbour_total": (C.c_int64, [vp]),
        "orc_get_neighbours": (C.c_int, [vp, _I64, _I64, _D]),
        "orc_dsph_prestep": (C.c_int, [vp, _D]),
        "orc_aero_velocity": (C.c_int, [vp]),
        "orc_detect_surface": (C.c_int, [vp]),
        "orc_dissipation_terms": (C.c_int, [vp]),
        "orc_particle_shift": (C.c_int, [vp]),
        "orc_boundary_update": (C.c_int, [vp]),
        "orc_acc_rrho": (C.c_int, [vp, C.c_double]),
        "orc_nb_iter": (C.c_int, [vp, C.c_double, _D]),
        "orc_find_timestep": (C.c_int, [vp, _D]),
        "orc_integrate_no_update": (C.c_int, [vp, C.POINTER(abi.StepInfo)]),
        "orc_integrate_step": (C.c_int, [vp, C.POINTER(abi.StepInfo)]),
        "orc_advance_frame": (C.c_int, [vp]),
        "orc_get_time": (C.c_int, [vp, _D, _D, _D]),
        "orc_set_time": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "orc_get_clock": (C.c_int, [vp, C.POINTER(abi.Clock)]),
        "orc_set_ipt": (C.c_int, [vp, C.POINTER(abi.IptSettings)]),
        "orc_moller": (C.c_int, [_D, _D, _D, _D, _D, _D, _D]),
        "orc_ipt_count": (C.c_int, [vp, _I64, _I64, _I64]),
        "orc_ipt_records": (C.c_int, [vp, C.c_void_p]),
        "orc_ipt_track": (C.c_int, [vp, C.c_int64, _D, _D, C.POINTER(C.c_int32), _I64, C.c_double, C.c_void_p]),
        "orc_get_inlet": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _I64, _I64]),
        "orc_set_clock": (C.c_int, [vp, C.POINTER(abi.Clock)]),
        "orc_qr_inverse": (C.c_int, [C.c_int, _D, _D]),
        "orc_min_eig": (C.c_double, [C.c_int, _D]),
        "orc_kernel": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_gradk_scalar": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "orc_set_mesh": (C.c_int, [vp, C.c_int64, _D, C.c_int64, _I64, _I64, _I64, C.c_int64, _I64, _I64,
                                   _D, _D, _D]),
        "orc_get_mesh_cells": (C.c_int, [vp, _D, _D, _D, _D]),
        "orc_mesh_lost": (C.c_int, [vp, _I64]),
        "orc_check_cell": (C.c_int, [vp, C.c_int64, _D]),
        "orc_mesh_lost_info": (C.c_int, [vp, _I64]),
        "orc_boundary_test": (C.c_int, [vp, _D, _I32]),
        "orc_first_cell_boundary": (C.c_int, [vp, _D, _I32]),
        "orc_line_intersection": (C.c_int, [vp, C.c_int64, _D, _D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIBS[kind] = lib
    return lib


def derive(settings):
    lib = load()
    d = abi.Derived()
    lib.orc_derive(C.byref(settings), C.byref(d))
    return d


def qr_inverse(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    dim = m.shape[0]
    out = np.zeros_like(m)
    ok = load().orc_qr_inverse(dim, _ptr(m), _ptr(out))
    return bool(ok), out


def min_eig(m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    return float(load().orc_min_eig(m.shape[0], _ptr(m)))


class Oracle:
    """One oracle simulation; method names follow the engine wrapper."""

    def __init__(self, case_or_settings, nthreads=None, kind="oracle"):
        self.lib = load(kind)
        settings = case_or_settings.settings if hasattr(case_or_settings, "settings") else case_or_settings
        self.settings = settings.copy()
        self.dim = int(settings.dim)
        if nthreads is None:
            nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        self.h = self.lib.orc_create(C.byref(self.settings), int(nthreads))
        if not self.h:
            raise RuntimeError("orc_create failed")
        if hasattr(case_or_settings, "x"):
            self.upload_case(case_or_settings)

    @property
    def n(self):
        nt, nb = np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.orc_get_counts(self.h, _ptr(nt, _I64), _ptr(nb, _I64)), "get_counts")
        return int(nt[0])

    def block_range(self, block):
        f, l = np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.orc_get_block_range(self.h, int(block), _ptr(f, _I64), _ptr(l, _I64)), "block_range")
        return int(f[0]), int(l[0])

    def set_block_planes(self, block, planes):
        self._planes = planes
        self._chk(self.lib.orc_set_block_planes(self.h, int(block), C.byref(planes)), "set_block_planes")

    def set_inlet(self, block, back, buffer):
        bk = np.ascontiguousarray(back, np.int64)
        bf = np.ascontiguousarray(buffer, np.int64).reshape(bk.size, -1)
        self._chk(self.lib.orc_set_inlet(self.h, int(block), bk.size, bf.shape[1], _ptr(bk, _I64), _ptr(bf, _I64)),
                  "set_inlet")

    def inlet(self, block):
        """(back, buffer) of the block's inlet columns, or None."""
        nc, nb = C.c_int(), C.c_int()
        self._chk(self.lib.orc_get_inlet(self.h, int(block), C.byref(nc), C.byref(nb), None, None), "get_inlet")
        if nc.value == 0:
            return None
        bk = np.zeros(nc.value, np.int64)
        bf = np.zeros((nc.value, nb.value), np.int64)
        self._chk(self.lib.orc_get_inlet(self.h, int(block), C.byref(nc), C.byref(nb), _ptr(bk, _I64), _ptr(bf, _I64)),
                  "get_inlet")
        return bk, bf

    @staticmethod
    def moller(v0, v1, v2, origin, direction, kind="oracle"):
        """(hit, dt, denom) of the reference's ray-triangle test."""
        lib = load(kind)
        a = [np.ascontiguousarray(q, np.float64) for q in (v0, v1, v2, origin, direction)]
        dt, dn = np.zeros(1), np.zeros(1)
        hit = lib.orc_moller(*[_ptr(q) for q in a], _ptr(dt), _ptr(dn))
        return bool(hit), float(dt[0]), float(dn[0])

    def set_ipt(self, settings):
        self._chk(self.lib.orc_set_ipt(self.h, C.byref(settings)), "set_ipt")

    def ipt_track(self, x, v, cell, part_id=None, t0=0.0):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        c = np.ascontiguousarray(cell, np.int32)
        pid = None if part_id is None else np.ascontiguousarray(part_id, np.int64)
        out = np.zeros(c.size, abi.IPT_DTYPE)
        self._chk(self.lib.orc_ipt_track(self.h, c.size, _ptr(x), _ptr(v), c.ctypes.data_as(C.POINTER(C.c_int32)),
                                         None if pid is None else _ptr(pid, _I64), float(t0),
                                         out.ctypes.data_as(C.c_void_p)), "ipt_track")
        return out

    def ipt_count(self):
        a, b, c = np.zeros(1, np.int64), np.zeros(1, np.int64), np.zeros(1, np.int64)
        self._chk(self.lib.orc_ipt_count(self.h, _ptr(a, _I64), _ptr(b, _I64), _ptr(c, _I64)), "ipt_count")
        return int(a[0]), int(b[0]), int(c[0])

    def ipt_records(self):
        out = np.zeros(self.ipt_count()[0], abi.IPT_DTYPE)
        self._chk(self.lib.orc_ipt_records(self.h, out.ctypes.data_as(C.c_void_p)), "ipt_records")
        return out

    def set_max_points(self, max_points):
        self._chk(self.lib.orc_set_max_points(self.h, int(max_points)), "set_max_points")

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what} failed with {rc}")

    def upload_case(self, case):
        self.upload_state(case.x, case.v, case.rho, case.p, case.m, case.b, case.n_bound)
        self.set_blocks(case.blocks)
        if getattr(case, "mesh", None) is not None:
            self.set_mesh(case.mesh)
        for blk, kw in (getattr(case, "planes", None) or {}).items():
            self.set_block_planes(blk, abi.make_planes(**kw))
        for blk, back, buffer in getattr(case, "inlets", None) or []:
            self.set_inlet(blk, back, buffer)
        if getattr(case, "max_points", None):
            self.set_max_points(case.max_points)

    def set_mesh(self, mesh):
        self._mesh = mesh
        a = lambda x, t=np.float64: np.ascontiguousarray(x, dtype=t)
        self._mesh_arrays = [a(mesh.verts), a(mesh.face_off, np.int64), a(mesh.face_verts, np.int64),
                             a(mesh.face_lr, np.int64), a(mesh.cell_off, np.int64), a(mesh.cell_faces, np.int64),
                             a(mesh.vert_vel), a(mesh.vert_p), a(mesh.vert_rho)]
        v, fo, fv, lr, co, cf, vv, vp, vr = self._mesh_arrays
        self._chk(self.lib.orc_set_mesh(self.h, mesh.nverts, _ptr(v), mesh.nfaces, _ptr(fo, _I64), _ptr(fv, _I64),
                                        _ptr(lr, _I64), mesh.ncells, _ptr(co, _I64), _ptr(cf, _I64), _ptr(vv),
                                        _ptr(vp), _ptr(vr)), "set_mesh")

    def mesh_cells(self):
        nc = self._mesh.ncells
        c, v = np.zeros((nc, self.dim)), np.zeros((nc, self.dim))
        p, r = np.zeros(nc), np.zeros(nc)
        self._chk(self.lib.orc_get_mesh_cells(self.h, _ptr(c), _ptr(v), _ptr(p), _ptr(r)), "get_mesh_cells")
        return c, v, p, r

    def mesh_lost(self):
        v = np.zeros(1, np.int64)
        self._chk(self.lib.orc_mesh_lost(self.h, _ptr(v, _I64)), "mesh_lost")
        return int(v[0])

    def check_cell(self, cell, point):
        p = np.zeros(3)
        p[:len(point)] = point
        return int(self.lib.orc_check_cell(self.h, int(cell), _ptr(p)))

    def lost_info(self):
        v = np.zeros(6, np.int64)
        self._chk(self.lib.orc_mesh_lost_info(self.h, _ptr(v, _I64)), "mesh_lost_info")
        return dict(zip(("lost", "deleted", "internal", "failed", "first_cell", "first_cell_deleted"),
                        (int(a) for a in v)))

    def boundary_test(self, point):
        """FindCell's boundary test of a point: (cross, wall, far)."""
        p = np.zeros(3)
        p[:len(point)] = point
        out = np.zeros(3, np.int32)
        self._chk(self.lib.orc_boundary_test(self.h, _ptr(p), _ptr(out, _I32)), "boundary_test")
        return int(out[0]), bool(out[1]), bool(out[2])

    def first_cell_boundary(self, x, v):
        """FirstCell's boundary branch for one point: (x, v after it, cross,
        delete)."""
        xv = np.zeros(6)
        xv[:len(x)] = x
        xv[3:3 + len(v)] = v
        out = np.zeros(2, np.int32)
        self._chk(self.lib.orc_first_cell_boundary(self.h, _ptr(xv), _ptr(out, _I32)), "first_cell_boundary")
        d = len(x)
        return xv[:d].copy(), xv[3:3 + d].copy(), int(out[0]), bool(out[1])

    def line_intersection(self, face, p1, p2):
        a, b = np.zeros(3), np.zeros(3)
        a[:len(p1)], b[:len(p2)] = p1, p2
        return int(self.lib.orc_line_intersection(self.h, int(face), _ptr(a), _ptr(b)))
    def upload_state(self, x, v, rho, p, m, b, n_bound):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        rho = np.ascontiguousarray(rho, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        m = np.ascontiguousarray(m, np.float64)
        b = np.ascontiguousarray(b, np.int32)
        self._chk(self.lib.orc_upload_state(self.h, x.shape[0], int(n_bound), _ptr(x), _ptr(v), _ptr(rho),
                                            _ptr(p), _ptr(m), _ptr(b, _I32)), "upload")

    def set_blocks(self, specs):
        arr = abi.make_blocks(specs)
        self._chk(self.lib.orc_set_blocks(self.h, len(specs), arr), "set_blocks")

    def set_serial_dissipation(self, on):
        self.lib.orc_set_serial_dissipation(self.h, int(on))

    def field(self, name, which=abi.STATE_NP1):
        w = abi.field_width(name, self.dim)
        n = self.n
        out = np.zeros(n * w)
        self._chk(self.lib.orc_download_field(self.h, which, abi.F[name], _ptr(out)), f"download {name}")
        return out.reshape(n, w) if w > 1 else out

    def set_field(self, name, values, which=abi.STATE_NP1):
        a = np.ascontiguousarray(values, np.float64).ravel()
        self._chk(self.lib.orc_upload_field(self.h, which, abi.F[name], _ptr(a)), f"upload {name}")

    def build_neighbours(self, brute=False):
        self._chk(self.lib.orc_build_neighbours(self.h, int(brute)), "neighbours")

    def neighbours(self):
        total = self.lib.orc_neighbour_total(self.h)
        off = np.zeros(self.n + 1, np.int64)
        idx = np.zeros(total, np.int64)
        d2 = np.zeros(total)
        self._chk(self.lib.orc_get_neighbours(self.h, _ptr(off, _I64), _ptr(idx, _I64), _ptr(d2)), "get_neighbours")
        return off, idx, d2

    def dsph_prestep(self):
        v = np.zeros(1)
        self._chk(self.lib.orc_dsph_prestep(self.h, _ptr(v)), "dsph")
        return float(v[0])

    def aero_velocity(self):
        self._chk(self.lib.orc_aero_velocity(self.h), "aero_velocity")

    def detect_surface(self):
        self._chk(self.lib.orc_detect_surface(self.h), "detect_surface")

    def dissipation_terms(self):
        self._chk(self.lib.orc_dissipation_terms(self.h), "dissipation")

    def particle_shift(self):
        self._chk(self.lib.orc_particle_shift(self.h), "shift")

    def boundary_update(self):
        self._chk(self.lib.orc_boundary_update(self.h), "boundary")

    def acc_rrho(self, npd):
        self._chk(self.lib.orc_acc_rrho(self.h, float(npd)), "acc_rrho")

    def nb_iter(self, npd):
        v = np.zeros(1)
        self._chk(self.lib.orc_nb_iter(self.h, float(npd), _ptr(v)), "nb_iter")
        return float(v[0])

    def find_timestep(self):
        v = np.zeros(1)
        self._chk(self.lib.orc_find_timestep(self.h, _ptr(v)), "find_timestep")
        return float(v[0])

    def integrate_no_update(self):
        info = abi.StepInfo()
        self._chk(self.lib.orc_integrate_no_update(self.h, C.byref(info)), "integrate_no_update")
        return info.as_dict()

    def integrate_step(self):
        info = abi.StepInfo()
        self._chk(self.lib.orc_integrate_step(self.h, C.byref(info)), "integrate")
        return info.as_dict()

    def advance_frame(self):
        self.lib.orc_advance_frame(self.h)

    def get_time(self):
        t, dt, cfl = np.zeros(1), np.zeros(1), np.zeros(1)
        self.lib.orc_get_time(self.h, _ptr(t), _ptr(dt), _ptr(cfl))
        return float(t[0]), float(dt[0]), float(cfl[0])

    def set_time(self, t, dt, cfl):
        self.lib.orc_set_time(self.h, float(t), float(dt), float(cfl))

    def clock(self):
        c = abi.Clock()
        self.lib.orc_get_clock(self.h, C.byref(c))
        return c

    def set_clock(self, clock):
        self.lib.orc_set_clock(self.h, C.byref(clock))
