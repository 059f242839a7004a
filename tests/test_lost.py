"""FindCell's handling of particles in no cell (oracle restatement,
src/Containment.cpp:734-805, src/Resid.cpp:481-519) against what the
geometry fixes, on the CPU:

- get_line_intersection is one-sided: a segment crossing an edge counts
  only when the cross product s x r of its direction s and the edge
  direction r is at least machine epsilon; checked against an independent
  segment-intersection test on random segments;
- on a structured quad mesh (far-field boundary edges, vertical edges
  running +y and horizontal ones +x), a point just above the top boundary
  crosses the top edge of the cell below it on the way to that cell's
  centre: odd parity and a far-field hit; the same with the top marked as a
  wall sets the wall flag instead; a point right of the mesh crosses only
  vertical edges leftwards, which the one-sided test misses: no hit;
- a point left of cell 0 hits cell 0's left edge once for cell 0 itself
  and once for every padded entry of the 500-nearest list (the reference's
  zero-initialised result vector), so the parity follows 1 + 500 - ncells;
- per lookup, particles in no cell above a far-field top are deleted (from
  both states, in order), beyond the right boundary their failure count
  rises by one per lookup and they are deleted at the lookup after it
  passes 10."""
import numpy as np
import pytest

from fjsph_amd import abi, cases


@pytest.fixture(scope="module")
def O():
    from oracle.pyoracle import Oracle
    return Oracle


def _settings(dx=1e-4):
    s = abi.default_settings(2)
    s.particle_step = dx
    s.aero_source = abi.MESH_INFL
    return s


def _mark(mesh, pred, marker):
    """Boundary edges whose midpoint satisfies pred get right cell `marker`."""
    lr = mesh.face_lr.reshape(-1, 2).copy()
    fv = mesh.face_verts.reshape(-1, 2)
    mid = 0.5 * (mesh.verts[fv[:, 0]] + mesh.verts[fv[:, 1]])
    sel = (lr[:, 1] < 0) & pred(mid)
    lr[sel, 1] = marker
    mesh.face_lr = lr
    return mesh


def _segments_cross(p, q, a, b):
    """Proper intersection of segments pq and ab (orientation signs)."""
    def orient(u, v, w):
        return np.sign((v[0] - u[0]) * (w[1] - u[1]) - (v[1] - u[1]) * (w[0] - u[0]))
    return orient(p, q, a) * orient(p, q, b) < 0 and orient(a, b, p) * orient(a, b, q) < 0


def test_line_intersection_is_one_sided(O):
    h = 0.01
    m = cases.quad_mesh(4, 3, h, [1.0, 0.0])
    o = O(_settings(), nthreads=1)
    o.set_mesh(m)
    fv = m.face_verts.reshape(-1, 2)
    rng = np.random.default_rng(5)
    counted = skipped = 0
    for _ in range(3000):
        f = int(rng.integers(0, fv.shape[0]))
        a, b = m.verts[fv[f, 0]], m.verts[fv[f, 1]]
        p1, p2 = rng.uniform(-0.01, 0.05, 2), rng.uniform(-0.01, 0.05, 2)
        s, r = p2 - p1, b - a
        cross = s[0] * r[1] - s[1] * r[0]
        if abs(cross) < 1e-9:
            continue
        t_num = (a - p1)
        # keep away from the segment ends, where rounding decides
        t = (t_num[0] * r[1] - t_num[1] * r[0]) / cross
        u = (t_num[0] * s[1] - t_num[1] * s[0]) / cross
        if min(abs(t), abs(t - 1), abs(u), abs(u - 1)) < 1e-9:
            continue
        expect = _segments_cross(p1, p2, a, b) and cross > 0
        got = o.line_intersection(f, p1, p2)
        assert got == int(expect), (f, p1, p2)
        if _segments_cross(p1, p2, a, b):
            counted += cross > 0
            skipped += cross < 0
    assert counted > 50 and skipped > 50


@pytest.mark.parametrize("marker,expect", [(-2, (1, False, True)), (-1, (1, True, False))])
def test_point_above_the_top(O, marker, expect):
    h = 0.01
    m = cases.quad_mesh(30, 20, h, [1.0, 0.0])  # 600 cells: no padding
    if marker == -1:
        m = _mark(m, lambda c: c[:, 1] > 20 * h - 1e-9, -1)
    o = O(_settings(), nthreads=1)
    o.set_mesh(m)
    for x in (0.0523, 0.1551, 0.2912):
        assert o.boundary_test([x, 20 * h + 0.0004]) == expect


def test_point_right_of_the_mesh_finds_no_crossing(O):
    h = 0.01
    m = cases.quad_mesh(30, 20, h, [1.0, 0.0])
    o = O(_settings(), nthreads=1)
    o.set_mesh(m)
    for y in (0.0123, 0.1047, 0.1862):
        assert o.boundary_test([30 * h + 0.0004, y]) == (0, False, False)


@pytest.mark.parametrize("nx,ny", [(20, 10), (15, 13), (30, 20)])
@pytest.mark.parametrize("marker", [-2, -1])
def test_padding_of_the_nearest_list(O, nx, ny, marker):
    h = 0.01
    m = cases.quad_mesh(nx, ny, h, [1.0, 0.0])
    m = _mark(m, lambda c: (c[:, 0] < 1e-9) & (c[:, 1] < h), marker)
    o = O(_settings(), nthreads=1)
    o.set_mesh(m)
    hits = 1 + max(0, 500 - nx * ny)
    cross, wall, far = o.boundary_test([-0.0001, 0.5 * h])
    assert cross == hits % 2
    assert (wall, far) == ((True, False) if marker == -1 else (False, True))


def _straddling(mesh_nx, mesh_ny, h, centre):
    case = cases.droplet_2d(diameter=4e-3, dx=2e-4, centre=centre, speed=3.0)
    case.settings.aero_source = abi.MESH_INFL
    case.mesh = cases.quad_mesh(mesh_nx, mesh_ny, h, [3.0, 0.0], p=case.settings.p_ref, rho=1.2)
    return case


def _lookup(o):
    o.build_neighbours()
    o.dsph_prestep()
    o.aero_velocity()


def test_far_field_deletion_per_lookup(O):
    case = _straddling(10, 5, 0.01, (0.05, 0.05))  # the droplet's top half is above the mesh
    o = O(case, nthreads=4)
    o.build_neighbours()
    o.dsph_prestep()
    x, lam = o.field("X"), o.field("LAM_NB")
    aero = lam <= case.settings.lam_cutoff
    gone = aero & (x[:, 1] > 0.05)
    keep_x = x[~gone]
    # the rays to the top-row centres that cross those cells' own top edges
    h = 0.01
    hits = np.zeros(case.n, int)
    for i in np.flatnonzero(gone):
        for c in range(10):
            a, b = np.array([c * h, 0.05]), np.array([(c + 1) * h, 0.05])
            hits[i] += _segments_cross(x[i], np.array([(c + 0.5) * h, 0.045]), a, b)
    assert np.all(hits[gone] >= 1)
    o.aero_velocity()
    info = o.lost_info()
    assert gone.sum() > 5 and o.n == case.n - gone.sum()
    assert info["lost"] == gone.sum() and info["deleted"] == gone.sum()
    assert info["failed"] == (hits[gone] % 2 == 0).sum() and info["internal"] == 0
    assert np.array_equal(o.field("X"), keep_x)
    assert np.array_equal(o.field("X", abi.STATE_N), keep_x)


def test_failure_count_then_deletion(O):
    case = _straddling(5, 10, 0.01, (0.05, 0.05))  # the droplet's right half is beyond the mesh
    o = O(case, nthreads=4)
    o.build_neighbours()
    o.dsph_prestep()
    x, lam = o.field("X"), o.field("LAM_NB")
    out = (lam <= case.settings.lam_cutoff) & (x[:, 0] > 0.05)
    n0 = case.n
    for k in range(1, 12):
        _lookup(o)
        nf = o.field("NFAILED")
        assert o.n == n0
        assert np.array_equal(nf[out], np.full(out.sum(), k)) and np.all(nf[~out] == 0)
        assert o.lost_info()["failed"] == k * out.sum()
    _lookup(o)
    assert o.n == n0 - out.sum() and o.lost_info()["deleted"] == out.sum()
