"""The 3D jet in crossflow on the GPU: a liquid jet issuing from a pipe with
ghost walls (Examples/Crossflow/pipe_3D.bmap) into a crossflow, with inlet
insertion, the pipe exit, deletion and the gas field of a mesh or its
surrogate.

Scaled-down parity against the FP64 oracle (counts, types, ids, inlet
columns exact; fields 1e-8 in FP64 and 1e-5 per step in mixed precision)
over steps that see insertions, deletions and mesh lookups; and a
full-size property run at about 1M particles (counts grow by whole
columns, no particle lost from the mesh, the surrogate within 1% of the
stream speed of the mesh lookup on the jet's aero particles).
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu


def _small(precision=abi.PREC_FP64, mode=0, mesh=False):
    dx = 3e-5
    c = cases.jet_in_crossflow_3d(n_target=0, diameter=3e-4, dx=dx, jet_layers=6, precision=precision, mode=mode,
                                  delete_above=5.6 * dx)
    if mesh:
        box = 0.5e-3
        cases.with_mesh(c, n_points=20000, kind="delaunay",
                        box=((-box, -0.6e-3, -box), (box, 0.6e-3, box)), speed=50.0)
    return c


def _pair(case):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    return Engine(case), Oracle(case, nthreads=8)


def _steps(eng, orc, n_steps, tol, check_each=True):
    eng.integrate_no_update()
    orc.integrate_no_update()
    ns = []
    for step in range(n_steps):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], f"step {step}: {ie} vs {io}"
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
        assert eng.n == orc.n, f"step {step}: {eng.n} vs {orc.n} particles"
        ns.append(eng.n)
        if check_each or step == n_steps - 1:
            assert np.array_equal(eng.field("B"), orc.field("B")), f"step {step}: particle types"
            assert np.array_equal(eng.field("PID"), orc.field("PID")), f"step {step}: particle ids"
            (eb, ef), (ob, of) = eng.inlet(1), orc.inlet(1)
            assert np.array_equal(eb, ob) and np.array_equal(ef, of), f"step {step}: inlet columns"
            for f in ("X", "V", "RHO", "P"):
                assert_close(f, eng.field(f), orc.field(f), tol)
    return ns


@pytest.mark.parametrize("mode", [0, 1], ids=["fixed", "dynamic"])
def test_jicf_parity_fp64(mode):
    case = _small(mode=mode)
    print(f"[jicf fp64] mode={mode} n={case.n} wall={case.n_bound}", flush=True)
    eng, orc = _pair(case)
    ns = _steps(eng, orc, 30, 1e-8, check_each=False)
    print("particle counts:", ns, flush=True)
    assert min(ns) < case.n, "no deletion seen"
    assert max(ns) > min(ns), "no insertion seen"
    # the ghost wall's densities and pressures follow the oracle too
    nb = case.n_bound
    for f in ("RHO", "P"):
        assert_close(f + "_wall", eng.field(f)[:nb], orc.field(f)[:nb], 1e-8)
    eng.close()


def test_jicf_parity_mesh_fp64():
    case = _small(mesh=True)
    print(f"[jicf mesh fp64] n={case.n} cells={case.mesh.ncells}", flush=True)
    eng, orc = _pair(case)
    ns = _steps(eng, orc, 20, 1e-8, check_each=False)
    cid = eng.field("CELLID")
    assert np.array_equal(cid, orc.field("CELLID"))
    assert (cid >= 0).sum() > 50
    assert eng.mesh_lost() == 0 and orc.mesh_lost() == 0
    for f in ("CELLV", "AF", "ACC"):
        assert_close(f, eng.field(f), orc.field(f), 1e-8)
    print("particle counts:", ns, flush=True)
    assert max(ns) > min(ns)
    eng.close()


def test_jicf_parity_mixed():
    case = _small(precision=abi.PREC_MIXED)
    print(f"[jicf mixed] n={case.n}", flush=True)
    eng, orc = _pair(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    ie, io = eng.integrate_step(), orc.integrate_step()
    assert ie["iterations"] == io["iterations"] or abs(ie["iterations"] - io["iterations"]) <= 1
    assert eng.n == orc.n
    for f in ("X", "V", "RHO", "P"):
        assert_close(f, eng.field(f), orc.field(f), 1e-5)
    eng.close()


def test_jicf_grows_at_1m_with_mesh_and_surrogate():
    """The configuration at full size: about 1M particles in a 0.2 m box
    meshed by a Delaunay triangulation (200k points here, to keep the test
    short; the bench uses 2M), potential flow past a cylinder at 50 m/s.
    Over the steps the count grows (column by column), no aero particle is in
    no cell, and the surrogate trained on the mesh lookup gives the jet's
    aero particles a gas velocity within 1% of 50 m/s (RMS) of the mesh
    lookup's."""
    from fjsph_amd import surrogate
    from fjsph_amd.engine import Engine
    case = cases.jet_in_crossflow_3d(n_target=1_000_000, precision=abi.PREC_MIXED)
    cases.with_mesh(case, n_points=200_000, kind="delaunay", box=((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1)), speed=50.0)
    ncol = case.n_columns
    print(f"[jicf 1M] n={case.n} columns={ncol} mesh cells={case.mesh.ncells}", flush=True)
    eng = Engine(case)
    eng.integrate_no_update()
    counts = [eng.n]
    for _ in range(40):
        eng.integrate_step()
        counts.append(eng.n)
    print("counts:", counts, flush=True)
    # each column inserts when its own back particle passes the insertion
    # plane; no deletion plane here, so the count only grows
    assert counts[-1] > counts[0]
    assert all(b >= a for a, b in zip(counts, counts[1:]))
    assert eng.mesh_lost() == 0
    cid = eng.field("CELLID")
    aero = cid >= 0
    x = eng.field("X")[aero]
    ref = np.concatenate([eng.field("CELLV")[aero], eng.field("CELLP")[aero, None],
                          eng.field("CELLRHO")[aero, None]], axis=1)
    eng.close()
    lab = Engine(case.settings)
    lab.set_mesh(case.mesh)
    sg = surrogate.train(case.mesh, lab, n_samples=1_000_000, steps=3000)
    lab.close()
    chk = Engine(case.settings)
    chk.set_surrogate(sg)
    pred = chk.surrogate_eval(x)
    chk.close()
    err = float(np.sqrt(np.mean(np.sum((pred[:, :3] - ref[:, :3]) ** 2, axis=1))))
    print(f"aero particles {aero.sum()}, surrogate velocity error {err:.3f} m/s", flush=True)
    assert aero.sum() > 1000
    assert err <= 0.01 * 50.0


# ------------------------------------------------------------- FirstCell

def _first_cell_case(wall):
    """The small jet with no free layers and a structured mesh whose lower
    boundary lies just above the aerodynamic entry plane (0.54 dx against
    0.5 dx): a particle leaving the pipe is below the mesh, in no cell, so
    FirstCell's boundary branch decides: across a far-field face it is
    deleted, across a wall face (`wall`: the lower faces marked -1) it is
    reflected onto the face.  The reference counts crossings only against
    each listed cell's own faces, on the way to that cell's centre, and
    stops when their number is even; a particle this close under the mesh
    crosses only the face above it, once, for this mesh spacing (4 dx) and
    these offsets (with a lower boundary at 1.5 dx, or a spacing of 2.5 dx,
    the reference, and the engine with it, stops with no net crossing)."""
    dx = 3e-5
    c = cases.jet_in_crossflow_3d(n_target=0, diameter=3e-4, dx=dx, jet_layers=0, precision=abi.PREC_FP64)
    h = 4.0 * dx
    lo = np.array([-5 * h + 0.5 * dx, 0.54 * dx, -5 * h + 1.7 * dx])
    hi = lo + np.array([10 * h, 6 * h, 10 * h])
    c.mesh = cases.kuhn_mesh(lo, hi, 3600, speed=50.0, rho_g=float(c.settings.rho_g), p_inf=float(c.settings.p_ref))
    c.settings.aero_source = abi.MESH_INFL
    if wall:
        m = c.mesh
        fv = m.face_verts.reshape(-1, 3)
        bottom = np.all(np.abs(m.verts[fv, 1] - lo[1]) < 1e-12, axis=1)
        lr = m.face_lr.reshape(-1, 2)
        lr[bottom, 1] = -1
        m.face_lr = np.ascontiguousarray(lr.reshape(-1))
    return c


@pytest.mark.parametrize("wall", [False, True], ids=["far_field", "wall"])
def test_first_cell_parity(wall):
    case = _first_cell_case(wall)
    print(f"[first cell] wall={wall} n={case.n} cells={case.mesh.ncells}", flush=True)
    eng, orc = _pair(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    seen = None
    for step in range(40):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], (step, ie, io)
        assert eng.n == orc.n, (step, eng.n, orc.n)
        le, lo_ = eng.lost_info(), orc.lost_info()
        assert le == lo_, (step, le, lo_)
        assert np.array_equal(eng.field("B"), orc.field("B"))
        assert np.array_equal(eng.field("PID"), orc.field("PID"))
        # a particle reflected onto the mesh's lower plane moves against the
        # column's flow: its neighbours' densities react strongly, and the
        # last bits of its position (the same in both up to FMA rounding of
        # the Newmark update) come back amplified: 1e-7 for the wall variant
        for f in ("X", "V", "RHO"):
            assert_close(f, eng.field(f), orc.field(f), 1e-7 if wall else 1e-8)
        if le["first_cell"] > 0:
            seen = (step, le)
            break
    assert seen is not None, "no particle left the pipe into no cell"
    step, info = seen
    print(f"  first FirstCell event at step {step}: {info}", flush=True)
    if wall:
        # reflected (the Newmark update then starts again from pn, as in
        # the reference), none deleted
        assert info["first_cell_deleted"] == 0 and eng.n == orc.n
    else:
        assert info["first_cell_deleted"] == info["first_cell"] > 0
    eng.close()
