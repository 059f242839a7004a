"""Restart through the engine (fjsph_restart_write / fjsph_restart_read): the
file holds the engine's state exactly, a fresh context loaded from it holds
the same state, and the run resumed from it matches the FP64 oracle resumed
from the same file (Read_HDF5 semantics: pn = pnp1 = the file's particles,
the clock's time, frame, CFL and CFL counters, last_frame_time = time,
src/H5IO.cpp:1215-1318; every other field starts from zero)."""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu

FP64_STEP = 1e-8
FIELDS = ("X", "V", "ACC", "RHO", "P", "RRHO", "M", "B", "PID", "CELLV", "CELLP", "CELLRHO", "CELLID")
IMAGE_OF = {"X": "x", "V": "v", "ACC": "acc", "RHO": "rho", "P": "p", "RRHO": "rrho", "M": "m", "B": "b",
            "PID": "pid", "CELLV": "cell_v", "CELLP": "cell_p", "CELLRHO": "cell_rho", "CELLID": "cell_id"}


def _oracle_from_image(img, settings):
    from oracle.pyoracle import Oracle
    specs = [dict(first=b.first, last=b.last, is_fluid=b.is_fluid, bound_solver=b.bound_solver, no_slip=b.no_slip,
                  vel=tuple(b.vel)) for b in img["blocks"]]
    case = cases.Case(name="restart", settings=settings, x=img["x"], v=img["v"], rho=img["rho"], p=img["p"],
                      m=img["m"], b=img["b"], n_bound=img["n_bound"], blocks=specs)
    orc = Oracle(case, nthreads=8)
    for k, b in enumerate(img["blocks"]):
        if b.is_fluid:
            orc.set_block_planes(k, img["planes"][k])
        if img["back"][k] is not None:
            orc.set_inlet(k, img["back"][k], np.asarray(img["buffer"][k]).reshape(-1, 4))
    for which in (abi.STATE_N, abi.STATE_NP1):
        for name in ("ACC", "RRHO", "CELLV", "CELLP", "CELLRHO", "PID", "CELLID"):
            orc.set_field(name, np.asarray(img[IMAGE_OF[name]], np.float64), which)
    ck = orc.clock()
    fc = img["clock"]
    for f in ("current_time", "cfl", "current_frame", "n_stable", "n_unstable", "next_part_id"):
        setattr(ck, f, getattr(fc, f))
    ck.last_frame_time = fc.current_time
    orc.set_clock(ck)
    return orc


@pytest.mark.parametrize("mode", [0, 1])
def test_restart_round_trip_and_resume(tmp_path, mode):
    from fjsph_amd import restart
    from fjsph_amd.engine import Engine

    case = cases.pipe_inlet_3d(mode=mode, n_buf=restart.NBUF)
    print(f"[restart] pipe inlet mode {mode}: n={case.n}", flush=True)
    eng = Engine(case)
    eng.integrate_no_update()
    for _ in range(3):  # one deletion happens in these steps
        eng.integrate_step()
    eng.advance_frame()
    path = str(tmp_path / "pipe_particles.h5")
    restart.write(eng, path, restart.default_meta(3, output_prefix="pipe"))

    # the file holds the engine's state exactly
    img = restart.read_host(path)
    st = {f: eng.field(f) for f in FIELDS}
    for f in FIELDS:
        assert np.array_equal(np.asarray(img[IMAGE_OF[f]], np.float64).reshape(st[f].shape), st[f]), f
    ck = eng.clock()
    fc = img["clock"]
    for f in ("current_time", "cfl", "current_frame", "n_stable", "n_unstable", "next_part_id"):
        assert getattr(fc, f) == getattr(ck, f), f
    assert fc.current_frame == 1 and fc.next_part_id == case.n
    back, buf = eng.inlet(0)
    assert np.array_equal(img["back"][0], back) and np.array_equal(np.asarray(img["buffer"][0]).reshape(buf.shape), buf)

    # a fresh context from the file holds the same state
    eng2 = Engine(case.settings)
    restart.read(eng2, path)
    assert eng2.counts() == eng.counts() and eng2.block_range(0) == eng.block_range(0)
    for f in FIELDS:
        assert np.array_equal(eng2.field(f), st[f]), f
        assert np.array_equal(eng2.field(f, abi.STATE_N), st[f]), f  # pn = pnp1
    b2, f2 = eng2.inlet(0)
    assert np.array_equal(b2, back) and np.array_equal(f2, buf)
    c2 = eng2.clock()
    assert c2.last_frame_time == ck.current_time
    for f in ("current_time", "cfl", "current_frame", "n_stable", "n_unstable", "next_part_id"):
        assert getattr(c2, f) == getattr(ck, f), f

    # the resumed run against the oracle resumed from the same file; the
    # steps see the next column advance and insertion (new particle ids
    # continue from the file's "Particle index to add")
    orc = _oracle_from_image(img, case.settings)
    ns = []
    for step in range(4):
        ie, io = eng2.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], f"step {step}: {ie} vs {io}"
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
        assert eng2.n == orc.n
        assert np.array_equal(eng2.field("B"), orc.field("B")), f"step {step}"
        assert np.array_equal(eng2.field("PID"), orc.field("PID")), f"step {step}"
        ns.append(eng2.n)
    for f in ("X", "V", "RHO", "P"):
        assert_close(f, eng2.field(f), orc.field(f), FP64_STEP)
    print("particle counts after the restart:", ns, flush=True)
    assert max(ns) > eng.n, "no insertion after the restart"
    assert eng2.field("PID").max() >= case.n, "new particles take ids from the counter"
