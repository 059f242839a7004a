"""Drift over long runs against the FP64 oracle (the stated bound: 1e-3
relative after 1000 steps):

- the plain 2D dam break in FP64 (1.0 x 2.0 m column in a 4.0 x 3.0 m tank,
  dx = 0.02 m, 6.5k particles), 1000 Newmark-Beta steps;
- a 3D droplet in crossflow (0.8 mm, dx = 0.05 mm, 2.1k particles) with the
  mixed-precision pair passes (FP32 pair terms, FP64 sums, positions and
  time) against the FP64 oracle: 500 steps with the adaptive time step (the
  droplet configuration's length), and 1000 steps at a fixed time step.
  Beyond about 700 adaptive steps the two runs no longer take the same time
  steps (a sub-iteration count that differs moves the CFL number), so the
  1000-step comparison holds the step fixed.
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


def test_droplet_drift_500_steps_mixed():
    case = cases.droplet_3d(diameter=0.8e-3, dx=5e-5, precision=abi.PREC_MIXED)
    print(f"[drift] droplet, mixed precision against the FP64 oracle: n={case.n}, 500 steps", flush=True)
    eng, orc, dts, its = _run(case, 500)
    te, to = eng.get_time()[0], orc.get_time()[0]
    print(f"  simulated time engine {te:.9e} oracle {to:.9e}; sub-iteration counts differ on "
          f"{(its[:, 0] != its[:, 1]).sum()} of 500 steps", flush=True)
    assert abs(te - to) <= 1e-9 * to
    rho0 = case.settings.rho_rest
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("P", eng.field("P"), orc.field("P"))):
        errs[f] = assert_close(f, g, r, DRIFT)
    print("  drift after 500 steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


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
