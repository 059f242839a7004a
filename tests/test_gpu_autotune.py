"""The tile-block autotuner (fjsph_autotune_blocks): it leaves the particle
state bit for bit as it was, the run that follows equals a fresh run given
the chosen setting, and a non-default tiling keeps FP64 parity with the
oracle (the tiling changes only the order of summation)."""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

FIELDS = ("X", "V", "ACC", "RHO", "P", "RRHO", "B", "PID", "CELLV", "CURVE", "NORM", "LAM_NB")


def _drops(precision):
    return cases.batched_droplets(n_drops=24, dx=5e-5, d_range=(0.3e-3, 0.8e-3), precision=precision)


def test_autotune_restores_state_and_matches_a_fresh_run():
    from fjsph_amd.engine import Engine
    case = _drops(abi.PREC_MIXED)
    print(f"[autotune] {case.n} particles in 24 droplets", flush=True)
    a, b = Engine(case), Engine(case)
    for e in (a, b):
        e.integrate_no_update()
        for _ in range(2):
            e.integrate_step()
    before = {f: a.field(f) for f in FIELDS}
    best, ms = a.autotune_blocks((128, 256, 512, 1024), reps=2)
    print("[autotune] ms per step's pair work:", {k: f"{v:.3f}" for k, v in ms.items()}, "best", best, flush=True)
    assert best in ms and a.block_targets() == best
    for f in FIELDS:
        assert np.array_equal(a.field(f), before[f]), f
    b.set_block_targets(best)
    for _ in range(2):
        ia, ib = a.integrate_step(), b.integrate_step()
        assert ia == ib
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f


@pytest.mark.parametrize("targets", [96, 1024])
def test_tilings_keep_fp64_parity(targets):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    case = _drops(abi.PREC_FP64)
    eng, orc = Engine(case), Oracle(case, nthreads=16)
    eng.set_block_targets(targets)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
    for f in ("X", "V", "RHO", "ACC"):
        assert_close(f, eng.field(f), orc.field(f), 1e-8)


@pytest.mark.parametrize("every", [1, 3])
def test_reorder_interval_keeps_fp64_parity(every):
    """The cell-order permutation every `every` steps (the builds in between
    index the particles through the cell order) changes only the order of
    summation."""
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    case = _drops(abi.PREC_FP64)
    eng, orc = Engine(case), Oracle(case, nthreads=16)
    eng.set_reorder_interval(every)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(4):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
    for f in ("X", "V", "RHO", "ACC"):
        assert_close(f, eng.field(f), orc.field(f), 1e-8)
    assert np.array_equal(eng.field("PID"), orc.field("PID"))


def test_autotune_on_a_dynamic_inlet_matches_a_fresh_run():
    """The autotuner's timing builds permute the particles; the restore puts
    every array back, and with it the inlet columns' back-particle slots, so
    a dynamic inlet (its buffer follows the back particle) steps on exactly
    as a fresh run with the chosen setting does, also on the steps between
    cell-order permutations (interval 4)."""
    from fjsph_amd.engine import Engine
    case = cases.pipe_inlet_3d(mode=1, precision=abi.PREC_MIXED)
    print(f"[autotune inlet] {case.n} particles, dynamic inlet", flush=True)
    a, b = Engine(case), Engine(case)
    for e in (a, b):
        e.set_reorder_interval(4)
        e.integrate_no_update()
        e.integrate_step()
    before = {f: a.field(f) for f in FIELDS}
    best, _ = a.autotune_blocks((64, 128, 256), reps=1)
    for f in FIELDS:
        assert np.array_equal(a.field(f), before[f]), f
    b.set_block_targets(best)
    for _ in range(6):
        ia, ib = a.integrate_step(), b.integrate_step()
        assert ia == ib
        assert a.n == b.n
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
