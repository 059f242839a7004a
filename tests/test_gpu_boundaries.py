"""Boundary solvers, moving walls and the Newmark rollback on the GPU against
the FP64 oracle (through the C-ABI).

Cases (all FP64, tolerance FP64_STEP normwise over whole steps, sub-iteration
counts and rollbacks equal):
  - DBC walls (src/Resid.cpp:21-76): the dam break and a 3D walled block;
  - ghost walls (src/Resid.cpp:78-141), free-slip and no-slip
    (Set_No_Slip, src/Resid.cpp:143-186), 2D and 3D, Newmark-Beta and RK4;
  - a wall moving by a velocity table (src/Newmark_Beta.cpp:69-84);
  - a ghost wall particle next to BUFFER fluid, whose density is clamped
    near the inlet (src/Newmark_Beta.cpp:163-178);
  - a step started at 8x the stable time step, which rolls back to pn and
    halves dt (src/Newmark_Beta.cpp:32-48).
The oracle's boundary and rollback code is pinned in tests/test_oracle_pins.py.
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

FP64_STEP = 1e-8


def _perturb(case, seed, dv=0.05, drho=0.005):
    rng = np.random.default_rng(seed)
    nb = case.n_bound
    case.v = case.v.copy()
    case.v[nb:] += rng.uniform(-dv, dv, size=case.v[nb:].shape)
    case.rho = case.rho.copy()
    case.rho[nb:] *= 1.0 + rng.uniform(-drho, drho, size=case.n - nb)
    s = case.settings
    B = s.rho_rest * s.speed_sound ** 2 / s.gam
    case.p = B * ((case.rho / s.rho_rest) ** s.gam - 1.0)
    return case


def _dam(solver, no_slip=0, seed=11):
    c = _perturb(cases.dam_break_2d(dx=0.05), seed)
    c.blocks[0]["bound_solver"] = solver
    c.blocks[0]["no_slip"] = no_slip
    return c


def _block3d(solver, no_slip=0, seed=12):
    c = cases.fluid_block(3, 10, dx=1e-3, jitter=0.02, seed=seed, with_walls=True, layers=3, H_fac=1.3)
    c.settings.speed_sound = 5.0
    c = _perturb(c, seed, dv=0.02, drho=0.002)
    c.blocks[0]["bound_solver"] = solver
    c.blocks[0]["no_slip"] = no_slip
    return c


def _pair(case):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    return Engine(case), Oracle(case, nthreads=8)


def _orc_motion(orc, block, t, v):
    import ctypes as C
    dp = C.POINTER(C.c_double)
    rc = orc.lib.orc_set_block_motion(orc.h, block, len(t), t.ctypes.data_as(dp), v.ctypes.data_as(dp))
    assert rc == 0


def _steps(eng, orc, n_steps, fields=("X", "V", "RHO", "P", "ACC", "RRHO")):
    ie, io = eng.integrate_no_update(), orc.integrate_no_update()
    assert ie["iterations"] == io["iterations"]
    for step in range(n_steps):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], f"step {step}: {ie} vs {io}"
        assert ie["rollbacks"] == io["rollbacks"], f"step {step}: {ie} vs {io}"
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
    for f in fields:
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)
    return ie, io


CASES = [
    ("dam_dbc", lambda: _dam(abi.DBC)),
    ("dam_ghost", lambda: _dam(abi.GHOST)),
    ("dam_ghost_noslip", lambda: _dam(abi.GHOST, no_slip=1)),
    ("dam_adami_noslip", lambda: _dam(abi.PRESSURE_G, no_slip=1)),
    ("block3d_dbc", lambda: _block3d(abi.DBC)),
    ("block3d_ghost_noslip", lambda: _block3d(abi.GHOST, no_slip=1)),
    ("block3d_adami_noslip", lambda: _block3d(abi.PRESSURE_G, no_slip=1)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_boundary_solver_steps(name, make):
    case = make()
    print(f"[boundary] {name}: n={case.n} nb={case.n_bound} dim={case.dim}", flush=True)
    eng, orc = _pair(case)
    nb = case.n_bound
    _steps(eng, orc, 4)
    # the wall fields the solver writes (density, pressure, velocity)
    for f in ("RHO", "P", "V"):
        assert_close(f + "_wall", eng.field(f)[:nb], orc.field(f)[:nb], FP64_STEP)
    if case.blocks[0]["bound_solver"] == abi.DBC:
        assert np.array_equal(eng.field("RHO")[:nb], case.rho[:nb])
    eng.close()


@pytest.mark.parametrize("name,make", [c for c in CASES if "ghost" in c[0] or "dbc" in c[0]],
                         ids=[c[0] for c in CASES if "ghost" in c[0] or "dbc" in c[0]])
def test_boundary_solver_steps_rk4(name, make):
    """Runge-Kutta 4 integrates DBC and ghost wall densities with the stage
    weights (src/Runge_Kutta.cpp)."""
    case = make()
    case.settings.integrator = abi.RUNGE_KUTTA
    print(f"[boundary rk4] {name}: n={case.n}", flush=True)
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
    for f in ("X", "V", "RHO", "P", "ACC", "RRHO"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)
    eng.close()


@pytest.mark.parametrize("solver", [abi.GHOST, abi.PRESSURE_G], ids=["ghost", "adami"])
def test_moving_wall(solver):
    """The tank's walls move with a velocity table: 0 before t = 0, then
    (0.4, 0) and from t = 1 ms (0.6, 0.2).  No-slip, so the wall velocity
    enters the fluid's viscous and pressure terms."""
    case = _dam(solver, no_slip=1, seed=13)
    case.settings.mu = 1e-3
    times = np.array([0.0, 1e-3])
    vels = np.array([[0.4, 0.0, 0.0], [0.6, 0.2, 0.0]])
    print(f"[moving wall] solver={solver}: n={case.n}", flush=True)
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    eng.set_block_motion(0, times, vels)
    _orc_motion(orc, 0, times, vels)
    _steps(eng, orc, 6)
    t, _, _ = orc.get_time()
    assert t > 1e-3, "the run must pass the second table time"
    vb = eng.field("V")[:case.n_bound]
    assert np.abs(vb).max() > 0.1
    eng.close()


def test_ghost_near_inlet_gpu():
    """A ghost wall particle inside BUFFER fluid (no FREE neighbour) takes the
    near-inlet clamp; next to FREE fluid it does not.  Engine = oracle."""
    s = abi.default_settings(2)
    s.particle_step = 1e-3
    s.H_fac = 1.3
    s.mu = 0.0
    s.sig = 0.0
    s.speed_sound = 5.0
    n_side = 9
    g = (np.arange(n_side) - (n_side - 1) / 2) * s.particle_step
    X, Y = np.meshgrid(g, g, indexing="xy")
    x = np.stack([X.ravel(), Y.ravel()], axis=1)
    centre = int(np.argmin(np.linalg.norm(x, axis=1)))
    order = np.r_[centre, np.delete(np.arange(len(x)), centre)]
    x = np.ascontiguousarray(x[order])
    v = np.ascontiguousarray(0.5 * x / np.abs(x).max())
    n = len(x)
    rho = np.full(n, 1000.0)
    m = rho * s.particle_step ** 2
    for kind in (abi.FREE, abi.BUFFER):
        b = np.full(n, kind, np.int32)
        b[0] = abi.BOUND
        case = cases.Case("ghost_inlet", s, x.copy(), v.copy(), rho.copy(), np.zeros(n), m.copy(), b, 1,
                          [dict(first=0, last=1, is_fluid=0, bound_solver=abi.GHOST),
                           dict(first=1, last=n, is_fluid=1)])
        from fjsph_amd.engine import Engine
        from oracle.pyoracle import Oracle
        eng, orc = Engine(case), Oracle(case, nthreads=2)
        eng.integrate_no_update()
        orc.integrate_no_update()
        re, ro = eng.field("RHO")[0], orc.field("RHO")[0]
        ge, go = eng.field("RRHO")[0], orc.field("RRHO")[0]
        assert abs(re - ro) <= 1e-12 * ro and abs(ge - go) <= 1e-9 * max(abs(go), 1.0), (kind, re, ro, ge, go)
        if kind == abi.BUFFER:
            assert re == 1000.0 and ge == 0.0
        else:
            assert re < 1000.0 and ge < 0.0
        eng.close()


def test_newmark_rollback_gpu():
    """Steps forced to 8x the stable time step (delta_t_min) diverge, roll
    back to pn and halve dt; the engine does the same as the oracle."""
    base = cases.dam_break_2d(dx=0.05)
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    probe = Oracle(base, nthreads=8)
    probe.integrate_no_update()
    for _ in range(3):
        probe.integrate_step()
    _, dt_stable, _ = probe.get_time()
    base.settings.delta_t_min = 8.0 * dt_stable
    print(f"[rollback] dam n={base.n} delta_t_min={base.settings.delta_t_min:.3e}", flush=True)
    eng, orc = Engine(base), Oracle(base, nthreads=8)
    eng.integrate_no_update()
    orc.integrate_no_update()
    total = 0
    for step in range(2):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["rollbacks"] == io["rollbacks"], (step, ie, io)
        assert ie["iterations"] == io["iterations"], (step, ie, io)
        assert ie["dt"] == io["dt"]
        total += ie["rollbacks"]
    assert total >= 1
    for f in ("X", "V", "RHO", "P", "ACC", "RRHO"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)
    eng.close()
