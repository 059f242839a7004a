"""NumPy restatement of the reference's TAU mesh and solution readers --
TEST INFRASTRUCTURE ONLY (tests/ import it; nothing in the engine does).

Reads NetCDF through scipy.io.netcdf_file (an independent reader of the
classic format) and assembles what the reference builds:

  Read_tau_mesh_EDGE + Place_Edges (src/CDFIO.cpp:830-1100), 2D
  Read_tau_mesh_FACE + Place_Faces (src/CDFIO.cpp:1103-1356), 3D
  Get_Coordinates (src/CDFIO.cpp:462-628): in 2D the missing one of
      points_xc/yc/zc is the dropped axis
  Read_SOLUTION (src/CDFIO.cpp:655-822): density, pressure, velocity
      components by offset axis (2D), through vertices_in_use (2D)

into the flat arrays of fjsph_set_mesh: verts, CSR faces, left/right per
face, CSR cells (each cell's faces in face order: the reference appends
them face by face under a critical section, so with one thread the order is
face order; the set is the same with any thread count), vertex solution.
"""
import numpy as np
from scipy.io import netcdf_file


def _read(path):
    f = netcdf_file(path, "r", mmap=False)
    dims = dict(f.dimensions)
    vars_ = {k: np.array(v.data) for k, v in f.variables.items()}
    f.close()
    return dims, vars_


def _coordinates(vars_, dim):
    names = ["points_xc", "points_yc", "points_zc"]
    present = [i for i, n in enumerate(names) if n in vars_]
    if dim == 3:
        assert present == [0, 1, 2]
        ignored = 0
    else:
        assert len(present) == 2
        ignored = [i for i in range(3) if i not in present][0] + 1
    return np.stack([vars_[names[i]].astype(np.float64) for i in present], axis=1), ignored


def _place(left, right, nelem):
    """cells' faces in face order; boundary faces are those with right < 0"""
    cell_faces = [[] for _ in range(nelem)]
    surf = []
    for i, (l, r) in enumerate(zip(left, right)):
        cell_faces[l].append(i)
        if r >= 0:
            cell_faces[r].append(i)
        else:
            surf.append(i)
    off = np.zeros(nelem + 1, np.int64)
    off[1:] = np.cumsum([len(c) for c in cell_faces])
    flat = np.array([f for c in cell_faces for f in c], np.int64)
    return off, flat, np.array(surf, np.int64)


def read_tau(mesh_path, sol_path, dim, offset_axis=2, scale=1.0):
    dims, mv = _read(mesh_path)
    if dim == 2:
        faces = mv["points_of_element_edges"].astype(np.int64)
        left = mv["left_element_of_edges"].astype(np.int64)
        right = mv["right_element_of_edges"].astype(np.int64)
        used = mv["vertices_in_use"].astype(np.int64) if "vertices_in_use" in mv else None
        face_list = [list(r) for r in faces]
    else:
        face_list = []
        if "points_of_triangles" in mv:
            face_list += [list(r) for r in mv["points_of_triangles"].astype(np.int64)]
        if "points_of_quadrilaterals" in mv:
            face_list += [list(r) for r in mv["points_of_quadrilaterals"].astype(np.int64)]
        left = mv["left_element_of_faces"].astype(np.int64)
        right = mv["right_element_of_faces"].astype(np.int64)
        used = None
    verts, ignored = _coordinates(mv, dim)
    verts = verts * scale
    nelem = dims["no_of_elements"]
    face_off = np.zeros(len(face_list) + 1, np.int64)
    face_off[1:] = np.cumsum([len(fc) for fc in face_list])
    face_verts = np.array([q for fc in face_list for q in fc], np.int64)
    cell_off, cell_faces, surf = _place(left, right, nelem)
    if "boundarymarker_of_surfaces" in mv:
        markers = mv["boundarymarker_of_surfaces"].astype(np.int64)
    else:
        markers = right[surf]
    out = dict(verts=verts, face_off=face_off, face_verts=face_verts,
               face_lr=np.stack([left, right], axis=1).ravel(), cell_off=cell_off, cell_faces=cell_faces,
               surf=surf, markers=markers, ignored=ignored)
    nv = verts.shape[0]
    if sol_path is None:
        out.update(vert_vel=np.zeros((nv, dim)), vert_p=np.zeros(nv), vert_rho=np.zeros(nv))
        return out
    _, sv = _read(sol_path)
    if dim == 3:
        comp = ["x_velocity", "y_velocity", "z_velocity"]
    else:
        comp = {1: ["y_velocity", "z_velocity"], 2: ["x_velocity", "z_velocity"],
                3: ["x_velocity", "y_velocity"]}[offset_axis]
    src = used if used is not None else np.arange(nv)
    out["vert_vel"] = np.stack([sv[c].astype(np.float64)[src] for c in comp], axis=1)
    out["vert_p"] = sv["pressure"].astype(np.float64)[src]
    out["vert_rho"] = sv["density"].astype(np.float64)[src]
    return out


def point_in_cell_2d(mesh, cell, p):
    """Even-odd ray test of a 2D cell (its edges), by exact rational-free
    float arithmetic along +x: the brute-force point location the tests use
    as the checker of the engine's cell search."""
    v = mesh["verts"]
    inside = False
    for k in range(mesh["cell_off"][cell], mesh["cell_off"][cell + 1]):
        f = mesh["cell_faces"][k]
        a, b = v[mesh["face_verts"][mesh["face_off"][f]]], v[mesh["face_verts"][mesh["face_off"][f] + 1]]
        if (a[1] > p[1]) != (b[1] > p[1]):
            xc = a[0] + (p[1] - a[1]) * (b[0] - a[0]) / (b[1] - a[1])
            if xc > p[0]:
                inside = not inside
    return inside
