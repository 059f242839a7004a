"""TAU mesh ingestion on the CPU (include/fjsph_tau.h): the engine library's
NetCDF reader and TAU assembly against the restatement in
oracle/tau_oracle.py, which reads through scipy's independent NetCDF
reader.  Data: the RAE2822 leading-edge piece (tests/golden, cut from the
reference's Examples/RAE2822 by tests/golden/make_rae2822_subset.py), the
full example when the reference tree is present, and a 3D face mesh with
record variables written here."""
import ctypes as C
import os
import subprocess
import textwrap

import numpy as np
import pytest
from scipy.io import netcdf_file

from fjsph_amd import abi, tau
from oracle import tau_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
LE_MESH = os.path.join(GOLD, "rae2822_le.grid.edges")
LE_SOL = os.path.join(GOLD, "rae2822_le.pval")
RAE = "/root/reference/Examples/RAE2822"


@pytest.fixture(scope="module")
def lib():
    try:
        return tau.load()
    except Exception as e:
        pytest.skip(str(e))


def _same(mesh, ref):
    assert np.array_equal(mesh.verts, ref["verts"])
    assert np.array_equal(mesh.face_off, ref["face_off"])
    assert np.array_equal(mesh.face_verts, ref["face_verts"])
    assert np.array_equal(mesh.face_lr.ravel(), ref["face_lr"])
    assert np.array_equal(mesh.cell_off, ref["cell_off"])
    assert np.array_equal(mesh.cell_faces, ref["cell_faces"])
    assert np.array_equal(mesh.vert_vel, ref["vert_vel"])
    assert np.array_equal(mesh.vert_p, ref["vert_p"])
    assert np.array_equal(mesh.vert_rho, ref["vert_rho"])


def test_leading_edge_piece_matches_oracle(lib):
    m = tau.TauMesh(LE_MESH, LE_SOL, dim=2, offset_axis=2, scale=1.0)
    mesh = m.mesh()
    ref = tau_oracle.read_tau(LE_MESH, LE_SOL, 2, offset_axis=2)
    _same(mesh, ref)
    z = m.sizes
    assert (z.dim, z.ignored_axis, z.ncells, z.nfaces, z.nverts) == (2, 2, 1866, 3830, 1965)
    assert z.n_sol_points == 2 * 1965
    faces, markers = m.surfaces()
    assert np.array_equal(faces, ref["surf"]) and np.array_equal(markers, ref["markers"])
    assert (markers == -1).sum() == 60  # the airfoil wall
    # every cell of the piece is a quadrilateral: four edges each
    assert np.all(np.diff(mesh.cell_off) == 4)
    # the mapping through vertices_in_use matters: the first plane in
    # mesh-vertex order is not the mesh's field
    f = netcdf_file(LE_SOL, "r", mmap=False)
    assert not np.array_equal(np.array(f.variables["pressure"].data[:z.nverts]), mesh.vert_p)
    f.close()
    m.close()


def test_scale_and_offset_axis(lib):
    ref = tau_oracle.read_tau(LE_MESH, LE_SOL, 2, offset_axis=3, scale=0.25)
    mesh = tau.read(LE_MESH, LE_SOL, dim=2, offset_axis=3, scale=0.25)
    _same(mesh, ref)
    # offset axis 3 takes (x, y) velocities: y is zero in this 2D solution
    assert np.all(mesh.vert_vel[:, 1] == 0.0) and np.any(mesh.vert_vel[:, 0] != 0.0)


@pytest.mark.skipif(not os.path.isdir(RAE), reason="reference example not present")
def test_full_rae2822_example(lib):
    mesh_path, sol_path = os.path.join(RAE, "mesh.grid.conf.edges"), os.path.join(RAE, "sol.pval.10000")
    m = tau.TauMesh(mesh_path, sol_path, dim=2, offset_axis=2)
    mesh = m.mesh()
    ref = tau_oracle.read_tau(mesh_path, sol_path, 2, offset_axis=2)
    _same(mesh, ref)
    z = m.sizes
    assert (z.ncells, z.nfaces, z.nverts, z.n_sol_points, z.nsurf) == (23552, 47504, 23952, 47904, 800)
    _, markers = m.surfaces()
    # the older converter's layout: right element -1 on the wall (no_of_wall_edges),
    # -2 on the far field (no_of_farfield_edges)
    assert (markers == -1).sum() == 304 and (markers == -2).sum() == 496
    m.close()


@pytest.mark.skipif(not os.path.isdir(RAE), reason="reference example not present")
def test_rae2822_bmap_angle_and_gravity(lib):
    s = abi.default_settings(2)
    s.grav[0], s.grav[1] = 0.0, -9.81
    a = tau.read_bmap(os.path.join(RAE, "mesh.grid.conf.bmap"), s)
    assert a == pytest.approx(np.deg2rad(12.0), rel=1e-15)
    assert (s.grav[0], s.grav[1]) == pytest.approx((9.81 * np.sin(a), -9.81 * np.cos(a)), rel=1e-15)


def test_bmap_gravity_rotation_3d(lib, tmp_path):
    p = tmp_path / "t.bmap"
    p.write_text("  Markers: 4\n # Angle alpha (degree): 45\n       Angle alpha (degree): 30.0\n block end\n")
    s = abi.default_settings(3)
    s.grav[0], s.grav[1], s.grav[2] = 0.0, 0.0, -9.81
    a = tau.read_bmap(str(p), s)
    assert a == pytest.approx(np.pi / 6, rel=1e-15)
    # Read_BMAP: grav.z scaled by cos(alpha); grav.x from the old y component (0 here)
    assert (s.grav[0], s.grav[1], s.grav[2]) == pytest.approx((0.0, 0.0, -9.81 * np.cos(a)), rel=1e-15)


def _pyramid_tet_files(tmp_path):
    """A pyramid (quad base) and a tetrahedron glued on one of its sides, in
    the 3D face layout, with the point coordinates and the solution as
    record variables (no_of_points unlimited)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.0], [1.6, 0.5, 0.5]], float)
    tris = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4], [1, 2, 5], [1, 4, 5], [2, 4, 5]], np.int32)
    quad = np.array([[0, 1, 2, 3]], np.int32)
    left = np.array([0, 0, 0, 0, 1, 1, 1, 0], np.int32)
    right = np.array([-1, 1, -2, -2, -2, -2, -2, -1], np.int32)
    mk = np.array([3, 4, 4, 4, 4, 4, 3], np.int32)
    mp = str(tmp_path / "pt.grid")
    f = netcdf_file(mp, "w", version=2)
    f.createDimension("no_of_points", None)
    for k, d in (("no_of_elements", 2), ("no_of_faces", 8), ("no_of_triangles", 7), ("points_per_triangle", 3),
                 ("no_of_quadrilaterals", 1), ("points_per_quadrilateral", 4), ("no_of_surfaceelements", 7)):
        f.createDimension(k, d)
    for k, a, dims in (("points_of_triangles", tris, ("no_of_triangles", "points_per_triangle")),
                       ("points_of_quadrilaterals", quad, ("no_of_quadrilaterals", "points_per_quadrilateral")),
                       ("left_element_of_faces", left, ("no_of_faces",)),
                       ("right_element_of_faces", right, ("no_of_faces",)),
                       ("boundarymarker_of_surfaces", mk, ("no_of_surfaceelements",))):
        f.createVariable(k, ">i4", dims)[:] = a
    for k, c in (("points_xc", 0), ("points_yc", 1), ("points_zc", 2)):
        f.createVariable(k, ">f8", ("no_of_points",))[:] = v[:, c]
    f.close()
    sp = str(tmp_path / "pt.pval")
    g = netcdf_file(sp, "w", version=1)
    g.createDimension("no_of_points", None)
    rng = np.random.default_rng(3)
    for k in ("x_velocity", "y_velocity", "z_velocity", "pressure", "density"):
        g.createVariable(k, ">f8", ("no_of_points",))[:] = rng.uniform(1.0, 2.0, 6)
    g.createVariable("cell_flag", ">i2", ("no_of_points",))[:] = np.arange(6, dtype=np.int16)
    g.close()
    return mp, sp


def test_face_mesh_with_record_variables(lib, tmp_path):
    mp, sp = _pyramid_tet_files(tmp_path)
    m = tau.TauMesh(mp, sp, dim=3)
    mesh = m.mesh()
    ref = tau_oracle.read_tau(mp, sp, 3)
    _same(mesh, ref)
    assert (m.sizes.ncells, m.sizes.nfaces, m.sizes.nsurf, m.sizes.ignored_axis) == (2, 8, 7, 0)
    # triangles first, then the quadrilateral
    assert list(np.diff(mesh.face_off)) == [3] * 7 + [4]
    faces, markers = m.surfaces()
    assert list(faces) == [0, 2, 3, 4, 5, 6, 7] and list(markers) == [3, 4, 4, 4, 4, 4, 3]
    # the pyramid has 5 faces, the tetrahedron 4
    assert list(np.diff(mesh.cell_off)) == [5, 4]
    m.close()


def test_reader_errors(lib, tmp_path):
    with pytest.raises(tau.TauError, match="cannot open"):
        tau.read(str(tmp_path / "none.grid"), None, dim=2)
    bad = tmp_path / "bad.grid"
    bad.write_bytes(b"HDF\x01" + b"\x00" * 60)
    with pytest.raises(tau.TauError, match="not a NetCDF"):
        tau.read(str(bad), None, dim=3)
    mp, sp = _pyramid_tet_files(tmp_path)
    with pytest.raises(tau.TauError, match="no_of_edges"):  # a 3D face file read as 2D
        tau.read(mp, None, dim=2)
    # a 2D edge file must lack exactly one coordinate (Get_Coordinates)
    tri = str(tmp_path / "tri.edges")
    f = netcdf_file(tri, "w", version=1)
    for k, d in (("no_of_elements", 1), ("no_of_edges", 3), ("points_per_edge", 2), ("no_of_points", 3)):
        f.createDimension(k, d)
    f.createVariable("points_of_element_edges", ">i4", ("no_of_edges", "points_per_edge"))[:] = [[0, 1], [1, 2], [2, 0]]
    f.createVariable("left_element_of_edges", ">i4", ("no_of_edges",))[:] = [0, 0, 0]
    f.createVariable("right_element_of_edges", ">i4", ("no_of_edges",))[:] = [-2, -2, -2]
    for k in ("points_xc", "points_yc", "points_zc"):
        f.createVariable(k, ">f8", ("no_of_points",))[:] = [0.0, 1.0, 0.0]
    f.close()
    with pytest.raises(tau.TauError, match="exactly two"):
        tau.read(tri, None, dim=2)


def test_library_exports_every_tau_symbol(lib, tmp_path):
    import re
    hdr = os.path.join(ROOT, "include", "fjsph_tau.h")
    text = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
    names = sorted(set(re.findall(r"\b(fjsph_[a-z0-9_]+)\s*\(", text)))
    assert len(names) == 8
    assert not [n for n in names if not hasattr(lib, n)]
    assert set(names) == set(tau.SIGNATURES)
    src = tmp_path / "sz.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include <stddef.h>
        #include "{hdr}"
        int main(void) {{ printf("%zu %zu\\n", sizeof(fjsph_tau_sizes), offsetof(fjsph_tau_sizes, n_sol_points)); return 0; }}
    """))
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(tmp_path / "sz"), str(src)], check=True)
    got = subprocess.run([str(tmp_path / "sz")], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in got] == [C.sizeof(tau.Sizes), tau.Sizes.n_sol_points.offset]
