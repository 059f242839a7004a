"""TAU mesh ingestion (ctypes over the calls of include/fjsph_tau.h in
fjsph_amd/libfjsph_engine.so): a TAU primary grid (face file in 3D, edge
file in 2D) and its solution become a cases.Mesh for Engine.set_mesh, or go
to an engine directly with set_mesh_tau.

    mesh = tau.read("mesh.grid.conf.edges", "sol.pval.10000", dim=2, offset_axis=2)
    eng.set_mesh(mesh)
"""
import ctypes as C

import numpy as np

from . import abi, cases, engine

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


class Sizes(C.Structure):
    _fields_ = [("dim", C.c_int32), ("ignored_axis", C.c_int32), ("nverts", C.c_int64), ("nfaces", C.c_int64),
                ("nface_verts", C.c_int64), ("ncells", C.c_int64), ("ncell_faces", C.c_int64),
                ("nsurf", C.c_int64), ("n_sol_points", C.c_int64)]


SIGNATURES = {
    "fjsph_tau_open": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_double, C.POINTER(C.c_void_p)]),
    "fjsph_tau_close": (None, [C.c_void_p]),
    "fjsph_tau_last_error": (C.c_char_p, []),
    "fjsph_tau_get_sizes": (C.c_int, [C.c_void_p, C.POINTER(Sizes)]),
    "fjsph_tau_get_arrays": (C.c_int, [C.c_void_p, _D, _I64, _I64, _I64, _I64, _I64, _D, _D, _D]),
    "fjsph_tau_get_surfaces": (C.c_int, [C.c_void_p, _I64, _I32]),
    "fjsph_set_mesh_tau": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fjsph_tau_read_bmap": (C.c_int, [C.c_char_p, C.POINTER(abi.Settings), _D]),
}
_LIB = None


class TauError(RuntimeError):
    pass


def load():
    global _LIB
    if _LIB is None:
        lib = engine.load()
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _chk(rc, what):
    if rc != abi.OK:
        raise TauError(f"{what}: status {rc}: {load().fjsph_tau_last_error().decode()}")


def _p(a, t):
    return a.ctypes.data_as(t)


class TauMesh:
    """An open TAU mesh (fjsph_tau_mesh); close() frees it."""

    def __init__(self, mesh_path, sol_path=None, dim=3, offset_axis=2, scale=1.0):
        lib = load()
        h = C.c_void_p()
        _chk(lib.fjsph_tau_open(str(mesh_path).encode(), None if sol_path is None else str(sol_path).encode(),
                                int(dim), int(offset_axis), float(scale), C.byref(h)), "tau_open")
        self.h = h
        self.sizes = Sizes()
        _chk(lib.fjsph_tau_get_sizes(self.h, C.byref(self.sizes)), "tau_get_sizes")

    def close(self):
        if self.h:
            load().fjsph_tau_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mesh(self):
        """The arrays as a cases.Mesh (the layout of fjsph_set_mesh)."""
        z = self.sizes
        d = int(z.dim)
        verts = np.zeros((z.nverts, d))
        fo = np.zeros(z.nfaces + 1, np.int64)
        fv = np.zeros(z.nface_verts, np.int64)
        lr = np.zeros((z.nfaces, 2), np.int64)
        co = np.zeros(z.ncells + 1, np.int64)
        cf = np.zeros(z.ncell_faces, np.int64)
        vel = np.zeros((z.nverts, d))
        p = np.zeros(z.nverts)
        rho = np.zeros(z.nverts)
        _chk(load().fjsph_tau_get_arrays(self.h, _p(verts, _D), _p(fo, _I64), _p(fv, _I64), _p(lr, _I64), _p(co, _I64),
                                         _p(cf, _I64), _p(vel, _D), _p(p, _D), _p(rho, _D)), "tau_get_arrays")
        return cases.Mesh(verts=verts, face_off=fo, face_verts=fv, face_lr=lr, cell_off=co, cell_faces=cf,
                          vert_vel=vel, vert_p=p, vert_rho=rho)

    def surfaces(self):
        n = int(self.sizes.nsurf)
        faces = np.zeros(n, np.int64)
        markers = np.zeros(n, np.int32)
        _chk(load().fjsph_tau_get_surfaces(self.h, _p(faces, _I64), _p(markers, _I32)), "tau_get_surfaces")
        return faces, markers

    def set_on(self, eng):
        """fjsph_set_mesh_tau: the mesh to an engine."""
        _chk(load().fjsph_set_mesh_tau(eng.h, self.h), "set_mesh_tau")
        eng._mesh = self.mesh()


def read(mesh_path, sol_path=None, dim=3, offset_axis=2, scale=1.0):
    """Read a TAU mesh and solution into a cases.Mesh."""
    m = TauMesh(mesh_path, sol_path, dim, offset_axis, scale)
    try:
        return m.mesh()
    finally:
        m.close()


def read_bmap(path, settings=None):
    """Angle of attack (radians) of a TAU boundary-mapping file; rotates
    settings.grav as Read_BMAP does when settings is given."""
    a = np.zeros(1)
    _chk(load().fjsph_tau_read_bmap(str(path).encode(), C.byref(settings) if settings is not None else None,
                                    _p(a, _D)), "tau_read_bmap")
    return float(a[0])
