"""TAU meshes on the GPU: the RAE2822 leading-edge piece (tests/golden, a
cut of the reference's Examples/RAE2822) and a 3D face mesh, read by the
engine library's TAU reader (fjsph_tau_open / fjsph_set_mesh_tau):
cell averages against the oracle, cell search against a brute-force point
location, a 2D droplet under the TAU gas field against the oracle, and the
surrogate trained on the real field."""
import os

import numpy as np
import pytest

from fjsph_amd import abi, cases, tau
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LE_MESH = os.path.join(HERE, "golden", "rae2822_le.grid.edges")
LE_SOL = os.path.join(HERE, "golden", "rae2822_le.pval")
FP64_STEP = 1e-8


def _cells_containing_2d(mesh, pts):
    """(P, C) boolean: point p inside cell c by the even-odd rule along +x
    over each cell's edges (vectorised brute force)."""
    v = mesh.verts
    fv = mesh.face_verts.reshape(-1, 2)
    a, b = v[fv[:, 0]], v[fv[:, 1]]
    px, py = pts[:, 0:1], pts[:, 1:2]
    straddle = (a[None, :, 1] > py) != (b[None, :, 1] > py)
    with np.errstate(divide="ignore", invalid="ignore"):
        xc = a[None, :, 0] + (py - a[None, :, 1]) * (b[None, :, 0] - a[None, :, 0]) / (b[None, :, 1] - a[None, :, 1])
    cross = (straddle & (xc > px)).astype(np.int64)
    lr = mesh.face_lr.reshape(-1, 2)
    ncell = mesh.cell_off.size - 1
    counts = np.zeros((pts.shape[0], ncell), np.int64)
    for side in (0, 1):
        ok = lr[:, side] >= 0
        np.add.at(counts.T, lr[ok, side], cross[:, ok].T)
    return counts % 2 == 1


def _engine_2d(case=None):
    from fjsph_amd.engine import Engine
    if case is None:
        s = abi.default_settings(2)
        s.particle_step = 2e-4
        s.aero_source = abi.MESH_INFL
        return Engine(s)
    return Engine(case)


def test_rae2822_piece_cells_and_lookup():
    m = tau.TauMesh(LE_MESH, LE_SOL, dim=2, offset_axis=2)
    mesh = m.mesh()
    print(f"[tau] RAE2822 piece: {mesh.ncells} cells, {mesh.nfaces} edges", flush=True)
    eng = _engine_2d()
    m.set_on(eng)
    from oracle.pyoracle import Oracle
    s = abi.default_settings(2)
    s.particle_step = 2e-4
    s.aero_source = abi.MESH_INFL
    orc = Oracle(s, nthreads=8)
    orc.set_mesh(mesh)
    for a, b in zip(eng.mesh_cells(), orc.mesh_cells()):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(11)
    lo, hi = mesh.verts.min(axis=0), mesh.verts.max(axis=0)
    pts = rng.uniform(lo, hi, size=(3000, 2))
    cell, vals = eng.mesh_eval(pts)
    inside = _cells_containing_2d(mesh, pts)
    found = cell >= 0
    assert found.sum() > 1000  # the piece is not a box: some points fall outside it or in the airfoil
    assert np.all(inside[np.nonzero(found)[0], cell[found]]), "a point outside the cell the engine gave"
    assert not np.any(inside[~found].any(axis=1)), "a point in a cell the engine did not find"
    _, cv, cp, cr = orc.mesh_cells()
    assert np.array_equal(vals[found, :2], cv[cell[found]]) and np.array_equal(vals[found, 2], cp[cell[found]])
    m.close()


def test_rae2822_droplet_steps_fp64():
    """A 2D droplet ahead of the leading edge in the RAE2822 gas field
    (mesh aero source, Gissler): cell ids, cell values and three steps
    against the oracle."""
    from oracle.pyoracle import Oracle
    case = cases.droplet_2d(diameter=4e-3, dx=2e-4, centre=(-0.015, 0.005))
    case.mesh = tau.read(LE_MESH, LE_SOL, dim=2, offset_axis=2)
    case.settings.aero_source = abi.MESH_INFL
    print(f"[tau] 2D droplet in the RAE2822 field: n={case.n}", flush=True)
    eng = _engine_2d(case)
    orc = Oracle(case, nthreads=8)
    for o in (eng, orc):
        o.build_neighbours()
        o.dsph_prestep()
        o.aero_velocity()
    cid = eng.field("CELLID")
    assert np.array_equal(cid, orc.field("CELLID")) and (cid >= 0).sum() > 50
    assert eng.mesh_lost() == 0 and orc.mesh_lost() == 0
    for f in ("CELLV", "CELLP", "CELLRHO"):
        assert np.array_equal(eng.field(f), orc.field(f)), f
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
    for f in ("X", "V", "RHO", "ACC", "AF"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)


def test_face_mesh_lookup_3d(tmp_path):
    from tests.test_tau import _pyramid_tet_files
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    mp, sp = _pyramid_tet_files(tmp_path)
    m = tau.TauMesh(mp, sp, dim=3)
    s = abi.default_settings(3)
    s.particle_step = 0.05
    s.aero_source = abi.MESH_INFL
    eng = Engine(s)
    m.set_on(eng)
    orc = Oracle(s, nthreads=2)
    orc.set_mesh(m.mesh())
    for a, b in zip(eng.mesh_cells(), orc.mesh_cells()):
        assert np.array_equal(a, b)
    pts = np.array([[0.5, 0.5, 0.3], [0.2, 0.5, 0.1], [1.025, 0.5, 0.375], [1.2, 0.5, 0.45], [2.0, 2.0, 2.0],
                    [0.5, 0.5, -0.1]])
    cell, _ = eng.mesh_eval(pts)
    assert list(cell) == [0, 0, 1, 1, -1, -1]
    m.close()


def test_surrogate_on_rae2822_piece():
    """The surrogate trained by the engine's kernels on the real TAU field
    (held-out gas-velocity error reported against the field's RMS speed)."""
    from fjsph_amd import surrogate
    m = tau.TauMesh(LE_MESH, LE_SOL, dim=2, offset_axis=2)
    eng = _engine_2d()
    m.set_on(eng)
    sg = surrogate.train(m.mesh(), eng, n_samples=400_000, steps=3000, batch=16384)
    x, y = surrogate.sample_mesh(m.mesh(), 50_000, seed=5, engine=eng)
    speed = float(np.sqrt(np.mean(np.sum(y[:, :2] ** 2, axis=1))))
    err = eng.surrogate_eval(x) - y
    vel_rms = float(np.sqrt(np.mean(np.sum(err[:, :2] ** 2, axis=1))))
    print(f"[tau] surrogate on the RAE2822 piece: loss {sg.loss[0]:.3e} -> {sg.loss[-1]:.3e}, gas velocity rms "
          f"error {vel_rms:.3f} m/s = {100 * vel_rms / speed:.2f}% of the field's rms speed {speed:.1f} m/s",
          flush=True)
    assert sg.loss[-100:].mean() < 0.05 * sg.loss[:10].mean()
    assert vel_rms < 0.05 * speed
    m.close()


def test_degenerate_cell_goes_to_the_overflow_list():
    """A cell with a vertex at the NetCDF fill value (9.97e36, as some
    points of the reference's RAE2822 example hold) must neither stretch
    the lookup grid nor fill it: it is kept apart and still found; the
    ordinary cells are found as before."""
    mesh = tau.read(LE_MESH, LE_SOL, dim=2, offset_axis=2)
    nv, nf, nc = mesh.nverts, mesh.nfaces, mesh.ncells
    far = 9.969209968386869e36
    verts = np.vstack([mesh.verts, [[0.2, 0.0], [0.3, 0.0], [0.25, far]]])
    tri = [(nv, nv + 1), (nv + 1, nv + 2), (nv + 2, nv)]
    deg = cases.Mesh(
        verts=verts, face_off=np.concatenate([mesh.face_off, mesh.face_off[-1] + 2 * np.arange(1, 4)]),
        face_verts=np.concatenate([mesh.face_verts, np.array(tri).ravel()]),
        face_lr=np.vstack([mesh.face_lr.reshape(-1, 2), [[nc, -2]] * 3]),
        cell_off=np.concatenate([mesh.cell_off, [mesh.cell_off[-1] + 3]]),
        cell_faces=np.concatenate([mesh.cell_faces, nf + np.arange(3)]),
        vert_vel=np.vstack([mesh.vert_vel, np.ones((3, 2))]), vert_p=np.concatenate([mesh.vert_p, [1.0] * 3]),
        vert_rho=np.concatenate([mesh.vert_rho, [1.0] * 3]))
    a, b = _engine_2d(), _engine_2d()
    a.set_mesh(mesh)
    b.set_mesh(deg)
    rng = np.random.default_rng(12)
    lo, hi = mesh.verts.min(axis=0), mesh.verts.max(axis=0)
    pts = rng.uniform(lo, hi, size=(2000, 2))
    ca, _ = a.mesh_eval(pts)
    cb, _ = b.mesh_eval(pts)
    assert np.array_equal(ca, cb)
    cd, vd = b.mesh_eval(np.array([[0.25, 1.0], [0.25, 1e30], [0.5, 1.0]]))
    assert list(cd) == [nc, nc, -1]
    assert np.array_equal(vd[0], [1.0, 1.0, 1.0, 1.0])
