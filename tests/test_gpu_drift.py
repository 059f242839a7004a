"""Drift over long runs against the FP64 oracle (the stated bound: 1e-3
relative after 1000 steps):

- the plain 2D dam break in FP64 (1.0 x 2.0 m column in a 4.0 x 3.0 m tank,
  dx = 0.02 m, 6.5k particles), 1000 Newmark-Beta steps;
- a 3D droplet in crossflow (0.8 mm, dx = 0.05 mm, 2.1k particles) with the
  mixed-precision pair passes (FP32 pair terms, FP64 sums, positions and
  time) against the FP64 oracle: 500 and 1000 steps with the adaptive time
  step (500 is the droplet configuration's length), and 1000 steps at a
  fixed time step.
- sub-iteration counts of the mixed path against the FP64 engine on the
  2D dam break at dx = 0.01 m (23k particles), 1000 steps: within one of
  each other on every step.  The first step starts from the converged
  state of the start-up pass and changes positions by ~1e-12 m; the FP64
  iterates settle on an exact fixed point (13 sub-iterations), which FP32
  pair terms cannot resolve, so the mixed path takes the force passes of
  such a step in FP64 once the rms change is within 100 ulps of the
  coordinates (before: the iteration limit, 21).
Errors are field-normwise (max |engine - oracle| / max |oracle|) and are
printed for every field."""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

DRIFT = 1e-3


def _run(case, steps, nthreads=16):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    eng, orc = Engine(case), Oracle(case, nthreads=nthreads)
    eng.integrate_no_update()
    orc.integrate_no_update()
    dts, its = [], []
    for k in range(steps):
        ie, io = eng.integrate_step(), orc.integrate_step()
        dts.append((ie["dt"], io["dt"]))
        its.append((ie["iterations"], io["iterations"]))
        if (k + 1) % 250 == 0:
            print(f"  step {k + 1}: time engine {eng.get_time()[0]:.9e} oracle {orc.get_time()[0]:.9e}", flush=True)
    dts, its = np.array(dts), np.array(its)
    return eng, orc, dts, its


def test_dam_break_drift_1000_steps_fp64():
    case = cases.dam_break_2d(dx=0.02)
    print(f"[drift] dam break FP64: n={case.n}, 1000 steps", flush=True)
    eng, orc, dts, its = _run(case, 1000)
    print(f"  sub-iteration counts differ on {(its[:, 0] != its[:, 1]).sum()} of 1000 steps; "
          f"max relative dt difference {np.max(np.abs(dts[:, 0] - dts[:, 1]) / dts[:, 1]):.2e}", flush=True)
    errs = {f: assert_close(f, eng.field(f), orc.field(f), DRIFT) for f in ("X", "V", "RHO", "P")}
    print("  drift after 1000 steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


@pytest.mark.parametrize("steps", [500, 1000])
def test_droplet_drift_adaptive_mixed(steps):
    case = cases.droplet_3d(diameter=0.8e-3, dx=5e-5, precision=abi.PREC_MIXED)
    print(f"[drift] droplet, mixed precision against the FP64 oracle: n={case.n}, {steps} steps", flush=True)
    eng, orc, dts, its = _run(case, steps)
    te, to = eng.get_time()[0], orc.get_time()[0]
    print(f"  simulated time engine {te:.9e} oracle {to:.9e}; sub-iteration counts differ on "
          f"{(its[:, 0] != its[:, 1]).sum()} of {steps} steps", flush=True)
    assert abs(te - to) <= 1e-9 * to
    rho0 = case.settings.rho_rest
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("P", eng.field("P"), orc.field("P"))):
        errs[f] = assert_close(f, g, r, DRIFT)
    print(f"  drift after {steps} steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


def test_droplet_drift_1000_steps_mixed_fixed_step():
    case = cases.droplet_3d(diameter=0.8e-3, dx=5e-5, precision=abi.PREC_MIXED)
    case.settings.delta_t_min = case.settings.delta_t_max = 3.2e-7  # the adaptive scheme's first steps
    print(f"[drift] droplet, mixed precision, fixed step 3.2e-7 s: n={case.n}, 1000 steps", flush=True)
    eng, orc, dts, its = _run(case, 1000)
    assert np.all(dts[:, 0] == dts[:, 1])
    print(f"  sub-iteration counts differ on {(its[:, 0] != its[:, 1]).sum()} of 1000 steps", flush=True)
    rho0 = case.settings.rho_rest
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("P", eng.field("P"), orc.field("P"))):
        errs[f] = assert_close(f, g, r, DRIFT)
    print("  drift after 1000 steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


def test_dam_break_mixed_iterations_1000_steps():
    """Sub-iteration counts of the mixed path within one of the FP64 engine's
    (itself pinned to the oracle by test_dam_break_drift_1000_steps_fp64) on
    every step, the time steps equal."""
    from fjsph_amd.engine import Engine
    runs = {}
    for prec in (abi.PREC_FP64, abi.PREC_MIXED):
        case = cases.dam_break_2d(dx=0.01)
        case.settings.precision = prec
        eng = Engine(case)
        eng.integrate_no_update()
        infos = [eng.integrate_step() for _ in range(1000)]
        eng.close()
        runs[prec] = (np.array([i["iterations"] for i in infos]), np.array([i["dt"] for i in infos]))
    d = runs[abi.PREC_MIXED][0] - runs[abi.PREC_FP64][0]
    print(f"[iterations] dam break dx=0.01, 1000 steps: FP64 mean {runs[abi.PREC_FP64][0].mean():.2f}, "
          f"mixed mean {runs[abi.PREC_MIXED][0].mean():.2f}, first step {runs[abi.PREC_FP64][0][0]} / "
          f"{runs[abi.PREC_MIXED][0][0]}, counts differ on {(d != 0).sum()} steps, max |diff| {np.abs(d).max()}",
          flush=True)
    assert np.abs(d).max() <= 1
    assert np.array_equal(runs[abi.PREC_MIXED][1], runs[abi.PREC_FP64][1])
