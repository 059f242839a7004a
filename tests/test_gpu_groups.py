"""Independent simulations in one context (fjsph_set_groups): a batch of
droplets of uneven size and speed stepped together on the GPU, each drop
against its OWN standalone run of the FP64 oracle.

Each drop must follow its standalone run: its own time step and CFL number
(src/Integration.cpp:233-303, 370-443), its own Newmark-Beta iteration count
and rollbacks (src/Newmark_Beta.cpp:10-52) and its own npd
(src/Shifting.cpp:16-68, 121).  Tolerances as in test_gpu_parity: FP64 1e-8
over steps (summation order only), mixed precision 1e-5 per step; iteration
counts, rollbacks and time steps equal.
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu


def _batch(n_drops, precision, seed=5, d_range=(0.3e-3, 0.7e-3)):
    return cases.batched_droplets(n_drops=n_drops, dx=5e-5, seed=seed, precision=precision, d_range=d_range)


def _run_batch(case, n_steps, tune=None):
    from fjsph_amd.engine import Engine
    eng = Engine(case)
    eng.set_groups(case.owner)
    if tune:
        eng.set_block_targets(tune)
    eng.integrate_no_update()
    infos = []
    for _ in range(n_steps):
        eng.integrate_step()
        infos.append(eng.group_info())
    fields = {f: eng.field(f) for f in ("X", "V", "RHO", "P", "ACC", "RRHO")}
    eng.close()
    return infos, fields


def _run_alone(case, g, n_steps):
    from oracle.pyoracle import Oracle
    sub = cases.drop_of_batch(case, g)
    orc = Oracle(sub, nthreads=8)
    orc.integrate_no_update()
    infos = [orc.integrate_step() for _ in range(n_steps)]
    fields = {f: orc.field(f) for f in ("X", "V", "RHO", "P", "ACC", "RRHO")}
    orc.close()
    return infos, fields


def _compare(case, n_steps, tol, tune=None):
    infos, fields = _run_batch(case, n_steps, tune)
    n_drops = int(case.owner.max()) + 1
    iters = set()
    for g in range(n_drops):
        sel = case.owner == g
        oinfos, ofields = _run_alone(case, g, n_steps)
        for step in range(n_steps):
            ie, io = infos[step][g], oinfos[step]
            assert ie["iterations"] == io["iterations"], (g, step, ie, io)
            assert ie["rollbacks"] == io["rollbacks"], (g, step, ie, io)
            assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"], (g, step, ie["dt"], io["dt"])
            iters.add(ie["iterations"])
        assert abs(infos[-1][g]["npd"] - oinfos[-1]["npd"]) <= tol * abs(oinfos[-1]["npd"])
        assert abs(infos[-1][g]["time"] - oinfos[-1]["time"]) <= 1e-10 * oinfos[-1]["time"]
        for f, ref in ofields.items():
            assert_close(f"{f}[drop {g}]", fields[f][sel], ref, tol)
    return infos, iters


def test_groups_match_standalone_fp64():
    case = _batch(6, abi.PREC_FP64)
    print(f"[groups fp64] {int(case.owner.max()) + 1} drops, n={case.n}", flush=True)
    infos, iters = _compare(case, 4, 1e-8)
    # the drops differ in size and speed: their time steps differ
    dts = [i["dt"] for i in infos[-1]]
    assert max(dts) > min(dts) * (1 + 1e-6)


def test_groups_match_standalone_mixed():
    case = _batch(5, abi.PREC_MIXED, seed=6)
    print(f"[groups mixed] {int(case.owner.max()) + 1} drops, n={case.n}", flush=True)
    _compare(case, 1, 1e-5)


def test_groups_independent_of_tile_blocks():
    """Other tile-block lengths give the same per-drop results (summation
    order only), with several drops in one row of cells."""
    case = _batch(4, abi.PREC_FP64, seed=7)
    print(f"[groups blocks] n={case.n}", flush=True)
    _compare(case, 2, 1e-8, tune=1024)


def rollback_batch():
    """Four drops at a time step forced to 1.5x the stable one
    (delta_t_min), drop 0 stirred by a random velocity of up to 1 m/s: drop
    0 diverges in the first step and rolls back, the others do not."""
    case = _batch(4, abi.PREC_FP64, seed=8)
    from oracle.pyoracle import Oracle
    probe = Oracle(cases.drop_of_batch(case, 0), nthreads=8)
    probe.integrate_no_update()
    probe.integrate_step()
    _, dt_stable, _ = probe.get_time()
    probe.close()
    sel = case.owner == 0
    case.v = case.v.copy()
    case.v[sel] += np.random.default_rng(3).uniform(-1.0, 1.0, size=case.v[sel].shape)
    case.settings.delta_t_min = 1.5 * dt_stable
    return case


def test_groups_rollback_alone():
    """A drop that diverges rolls back on its own (its own pn, its own
    halved time step, its lists built again while the others keep theirs
    and go on iterating), as its standalone run does; the others follow
    their standalone runs, which do not roll back."""
    case = rollback_batch()
    print(f"[groups rollback] n={case.n} delta_t_min={case.settings.delta_t_min:.3e}", flush=True)
    infos, _ = _compare(case, 2, 1e-8)
    rb = [i["rollbacks"] for i in infos[0]]
    assert rb[0] == 1 and sum(rb[1:]) == 0, rb
