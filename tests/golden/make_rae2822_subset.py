"""Write the RAE2822 leading-edge fixture: a piece of the reference's 2D TAU
example (Examples/RAE2822: mesh.grid.conf.edges, sol.pval.10000), in the
same NetCDF layout, small enough to travel with the tests.

Cells whose centre lies in x in (-0.03, 0.1), z in (-0.05, 0.06) (the
airfoil's leading edge, wall included) are kept with every edge that
touches them, in the original edge order.  Vertices and cells are
renumbered in ascending original order.  An edge that leaves the piece gets
the far-field marker -2 as its right element (with its kept cell on the
left); wall edges keep -1.  The solution file holds two planes of points
like the original (2 x the mesh points); the mesh's vertices_in_use sends
vertex k to a shuffled position of the first plane, so a reader that skips
the mapping reads wrong values.

Usage: python tests/golden/make_rae2822_subset.py [/root/reference]
"""
import os
import sys

import numpy as np
from scipy.io import netcdf_file

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "Examples", "RAE2822")
OUT = os.path.dirname(os.path.abspath(__file__))
BOX = (-0.03, 0.1, -0.05, 0.06)


def main():
    f = netcdf_file(os.path.join(SRC, "mesh.grid.conf.edges"), "r", mmap=False)
    x = f.variables["points_xc"].data.astype(np.float64)
    z = f.variables["points_zc"].data.astype(np.float64)
    e = f.variables["points_of_element_edges"].data.astype(np.int64)
    left = f.variables["left_element_of_edges"].data.astype(np.int64)
    right = f.variables["right_element_of_edges"].data.astype(np.int64)
    used = f.variables["vertices_in_use"].data.astype(np.int64)
    nelem = f.dimensions["no_of_elements"]
    f.close()
    s = netcdf_file(os.path.join(SRC, "sol.pval.10000"), "r", mmap=False)
    sol = {k: s.variables[k].data.astype(np.float64) for k in ("density", "x_velocity", "y_velocity", "z_velocity",
                                                                 "pressure")}
    nsol = s.dimensions["no_of_points"]
    s.close()

    # cell centres (vertex means over the edges, each vertex counted per edge)
    cx = np.zeros(nelem)
    cz = np.zeros(nelem)
    cnt = np.zeros(nelem)
    for side in (left, right):
        ok = side >= 0
        for k in (0, 1):
            np.add.at(cx, side[ok], x[e[ok, k]])
            np.add.at(cz, side[ok], z[e[ok, k]])
            np.add.at(cnt, side[ok], 1)
    cx /= cnt
    cz /= cnt
    keep = (cx > BOX[0]) & (cx < BOX[1]) & (cz > BOX[2]) & (cz < BOX[3])
    cells = np.nonzero(keep)[0]
    new_cell = np.full(nelem, -1, np.int64)
    new_cell[cells] = np.arange(cells.size)
    inl = (left >= 0) & keep[np.maximum(left, 0)]
    inr = (right >= 0) & keep[np.maximum(right, 0)]
    edges = np.nonzero(inl | inr)[0]
    verts = np.unique(e[edges].ravel())
    new_vert = np.full(x.size, -1, np.int64)
    new_vert[verts] = np.arange(verts.size)

    el, er = [], []
    for i in edges:
        li, ri = inl[i], inr[i]
        if li and ri:
            el.append(new_cell[left[i]]), er.append(new_cell[right[i]])
        elif li:
            el.append(new_cell[left[i]]), er.append(right[i] if right[i] < 0 else -2)
        else:
            el.append(new_cell[right[i]]), er.append(-2)
    nv, ne = verts.size, edges.size

    rng = np.random.default_rng(2822)
    slot = rng.permutation(nv)  # first-plane position of each vertex
    out_sol = {}
    for k, a in sol.items():
        o = np.zeros(2 * nv)
        o[slot] = a[used[verts]]
        o[nv:] = a[np.minimum(used[verts] + nsol // 2, nsol - 1)]
        out_sol[k] = o

    m = netcdf_file(os.path.join(OUT, "rae2822_le.grid.edges"), "w", version=1)
    m.createDimension("no_of_elements", cells.size)
    m.createDimension("no_of_edges", ne)
    m.createDimension("points_per_edge", 2)
    m.createDimension("no_of_points", nv)
    v = m.createVariable("points_of_element_edges", ">i4", ("no_of_edges", "points_per_edge"))
    v[:] = new_vert[e[edges]].astype(np.int32)
    v = m.createVariable("left_element_of_edges", ">i4", ("no_of_edges",))
    v[:] = np.array(el, np.int32)
    v = m.createVariable("right_element_of_edges", ">i4", ("no_of_edges",))
    v[:] = np.array(er, np.int32)
    v = m.createVariable("vertices_in_use", ">i4", ("no_of_points",))
    v[:] = slot.astype(np.int32)
    v = m.createVariable("points_xc", ">f8", ("no_of_points",))
    v[:] = x[verts]
    v = m.createVariable("points_zc", ">f8", ("no_of_points",))
    v[:] = z[verts]
    m.close()

    p = netcdf_file(os.path.join(OUT, "rae2822_le.pval"), "w", version=2)
    p.createDimension("no_of_points", 2 * nv)
    for k, a in out_sol.items():
        v = p.createVariable(k, ">f8", ("no_of_points",))
        v[:] = a
    p.close()
    print(f"{cells.size} cells, {ne} edges, {nv} vertices, {int((np.array(er) == -1).sum())} wall edges")


if __name__ == "__main__":
    main()
