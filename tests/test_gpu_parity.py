"""Engine (HIP, through the C-ABI) versus the FP64 oracle, pass by pass and
over whole steps.

Tolerances (field-normwise: max_i |engine_i - oracle_i| / max_i |oracle_i|):
  FP64 path    1e-9 per pass, 1e-8 over a few steps (summation order of the
               neighbour loops differs from the oracle's index order);
  mixed path   1e-5 per step (FP32 pair terms, FP64 accumulation), as the
               target states; 1e-3 drift over the multi-step runs; 1e-3 for
               the intermediate ALE shift velocity (see MIXED_VPERT).
Neighbour sets must match exactly.
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close, same_neighbour_sets

pytestmark = pytest.mark.gpu

FP64_PASS = 1e-9
FP64_STEP = 1e-8
MIXED_STEP = 1e-5
# The interior ALE shift velocity is the residual of a pair sum that cancels
# to about 1e-3 of its terms on a regular lattice: FP32 pair terms (relative
# error ~1e-7 each, ~250 per particle) leave ~1e-4 of it (6e-5 measured on
# the small jet; the FP64 path: 2e-13).  It moves positions by dt * vPert,
# so X stays within MIXED_STEP over a step (test_step_parity_mixed).
MIXED_VPERT = 1e-3


def _engine(case):
    from fjsph_amd.engine import Engine
    return Engine(case)


def _oracle(case):
    from oracle.pyoracle import Oracle
    return Oracle(case, nthreads=8)


def _perturb(case, seed=5, dv=0.05, drho=0.005):
    """Give the lattice a random velocity and density field so every pair
    term is exercised; p follows the Cole EOS."""
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


def _small_cases():
    dam = _perturb(cases.dam_break_2d(dx=0.05), seed=1)
    drop = _perturb(cases.droplet_3d(diameter=1.0e-3, dx=5e-5, precision=abi.PREC_FP64), seed=2)
    jet = _perturb(cases.jet_column_3d(n_target=6000, diameter=0.6e-3, dx=5e-5, precision=abi.PREC_FP64), seed=3)
    # Hydrostatic column, H = 2 dx: 16-cell tile blocks whose waves mix wall
    # lanes (no surface, shift or force terms of their own) with fluid lanes
    # that have fewer rows than the wall lanes in some tiles
    column = _perturb(cases.standing_column_2d(dx=0.02), seed=4)
    return [("dam2d", dam), ("drop3d", drop), ("jet3d_ale", jet), ("column2d", column)]


@pytest.mark.parametrize("name,case", _small_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_pass_parity_fp64(name, case):
    print(f"[pass parity] {name}: n={case.n} nb={case.n_bound} dim={case.dim}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    nb = case.n_bound
    fluid = slice(nb, case.n)
    eng.build_neighbours()
    orc.build_neighbours()
    off_e, idx_e = eng.neighbours()
    off_o, idx_o, _ = orc.neighbours()
    assert same_neighbour_sets(off_e, idx_e, off_o, idx_o), "neighbour sets differ"

    npd_e, npd_o = eng.dsph_prestep(), orc.dsph_prestep()
    assert abs(npd_e - npd_o) <= 1e-12 * abs(npd_o)
    for f in ("L", "GRADRHO", "NORM", "LAM", "LAM_NB", "COLOURG", "KERNSUM", "COLOUR"):
        assert_close(f, eng.field(f), orc.field(f), FP64_PASS)

    eng.aero_velocity()
    orc.aero_velocity()
    assert np.array_equal(eng.field("CELLID"), orc.field("CELLID"))
    assert_close("CELLV", eng.field("CELLV"), orc.field("CELLV"), 0.0)

    eng.detect_surface()
    orc.detect_surface()
    assert np.array_equal(eng.field("SURF")[fluid], orc.field("SURF")[fluid]), "surface flags differ"
    for f in ("NORM", "CURVE", "WOCCL", "NORM_CURVE"):
        assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], FP64_PASS)
    if case.settings.ale:
        assert np.array_equal(eng.field("SURFZONE"), orc.field("SURFZONE"))

    eng.dissipation_terms()
    orc.dissipation_terms()
    for f in ("AVISC", "DELTAD"):
        assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], FP64_PASS)

    if case.settings.ale:
        eng.particle_shift()
        orc.particle_shift()
        assert_close("VPERT", eng.field("VPERT")[fluid], orc.field("VPERT")[fluid], FP64_PASS)

    dt_e, dt_o = eng.find_timestep(), orc.find_timestep()
    assert abs(dt_e - dt_o) <= 1e-12 * dt_o
    eng.set_time(0.0, dt_o, case.settings.cfl)
    orc.set_time(0.0, dt_o, case.settings.cfl)

    if nb:
        eng.boundary_update()
        orc.boundary_update()
        for f in ("P", "RHO"):
            assert_close(f + "_bound", eng.field(f)[:nb], orc.field(f)[:nb], FP64_PASS)

    eng.acc_rrho(npd_o)
    orc.acc_rrho(npd_o)
    for f in ("ACC", "RRHO", "AF"):
        assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], FP64_PASS)

    le_e, le_o = eng.nb_iter(npd_o), orc.nb_iter(npd_o)
    assert abs(le_e - le_o) <= 1e-9 * abs(le_o)
    for f in ("X", "V", "RHO", "P"):
        assert_close(f, eng.field(f), orc.field(f), FP64_PASS)


@pytest.mark.parametrize("name,case", _small_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_step_parity_fp64(name, case):
    print(f"[step parity] {name}: n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    ie, io = eng.integrate_no_update(), orc.integrate_no_update()
    assert ie["iterations"] == io["iterations"]
    for step in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], f"step {step}: sub-iterations {ie} vs {io}"
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
        # rms_error = log10 of a difference of nearly equal iterates: its
        # conditioning is |x| / |dx| ~ 1e11, so compare in log units.
        assert abs(ie["rms_error"] - io["rms_error"]) <= 1e-3
    for f in ("X", "V", "RHO", "P", "ACC", "RRHO"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)


def _rk(case):
    case.settings.integrator = abi.RUNGE_KUTTA
    return case


@pytest.mark.parametrize("name,case", [(n, _rk(c)) for n, c in _small_cases() if n != "drop3d"],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_step_parity_rk4_fp64(name, case):
    """Runge-Kutta 4 (src/Runge_Kutta.cpp) on the GPU against the oracle."""
    print(f"[rk4 step parity] {name}: n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    ie, io = eng.integrate_no_update(), orc.integrate_no_update()
    assert abs(ie["rms_error"] - io["rms_error"]) <= 1e-3
    for step in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
        assert abs(ie["rms_error"] - io["rms_error"]) <= 1e-3
    for f in ("X", "V", "RHO", "P", "ACC", "RRHO"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)


def test_mesh_aero_fp64():
    """Mesh aero source (FindCell + Gissler) on a 3D droplet in a Delaunay
    mesh: cell ids, cell values and two full steps against the oracle."""
    case = cases.with_mesh(_perturb(cases.droplet_3d(diameter=1.0e-3, dx=5e-5, precision=abi.PREC_FP64), seed=6),
                           n_points=20000)
    print(f"[mesh aero] droplet n={case.n}, mesh cells={case.mesh.ncells}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    ce, co = eng.mesh_cells(), orc.mesh_cells()
    for a, b in zip(ce, co):
        assert np.array_equal(a, b)
    for o in (eng, orc):
        o.build_neighbours()
        o.dsph_prestep()
        o.aero_velocity()
    assert eng.mesh_lost() == 0 and orc.mesh_lost() == 0
    cid = eng.field("CELLID")
    assert np.array_equal(cid, orc.field("CELLID"))
    assert (cid >= 0).sum() > 100
    for f in ("CELLV", "CELLP", "CELLRHO"):
        assert np.array_equal(eng.field(f), orc.field(f)), f
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(2):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
    for f in ("X", "V", "RHO", "ACC", "AF"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)


def test_batched_droplets_fp64():
    """Many independent droplets of uneven size in one particle set (the
    batched-droplet configuration, scaled down): steps against the oracle."""
    case = cases.batched_droplets(n_drops=12, dx=5e-5, d_range=(0.2e-3, 0.5e-3), precision=abi.PREC_FP64)
    print(f"[batched droplets] n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
    for f in ("X", "V", "RHO", "ACC"):
        assert_close(f, eng.field(f), orc.field(f), FP64_STEP)


def _mixed(case):
    case.settings.precision = abi.PREC_MIXED
    return case


@pytest.mark.parametrize("name,case", [(n, _mixed(c)) for n, c in _small_cases()],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_pass_parity_mixed(name, case):
    """Mixed path pass by pass against the FP64 oracle: fields within
    MIXED_STEP normwise, flags (surface, surface zone, aero cell) exact."""
    print(f"[mixed pass parity] {name}: n={case.n} nb={case.n_bound} dim={case.dim}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    nb = case.n_bound
    fluid = slice(nb, case.n)
    eng.build_neighbours()
    orc.build_neighbours()
    off_e, idx_e = eng.neighbours()
    off_o, idx_o, _ = orc.neighbours()
    assert same_neighbour_sets(off_e, idx_e, off_o, idx_o), "neighbour sets differ"

    npd_e, npd_o = eng.dsph_prestep(), orc.dsph_prestep()
    assert abs(npd_e - npd_o) <= MIXED_STEP * abs(npd_o)
    errs = {}
    for f in ("L", "GRADRHO", "NORM", "LAM", "LAM_NB", "COLOURG", "KERNSUM", "COLOUR"):
        errs[f] = assert_close(f, eng.field(f), orc.field(f), MIXED_STEP)

    eng.aero_velocity()
    orc.aero_velocity()
    assert np.array_equal(eng.field("CELLID"), orc.field("CELLID"))

    eng.detect_surface()
    orc.detect_surface()
    assert np.array_equal(eng.field("SURF")[fluid], orc.field("SURF")[fluid]), "surface flags differ"
    for f in ("NORM", "CURVE", "WOCCL"):
        errs[f + "_surf"] = assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], MIXED_STEP)
    if case.settings.ale:
        assert np.array_equal(eng.field("SURFZONE"), orc.field("SURFZONE"))

    eng.dissipation_terms()
    orc.dissipation_terms()
    for f in ("AVISC", "DELTAD"):
        errs[f] = assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], MIXED_STEP)

    if case.settings.ale:
        eng.particle_shift()
        orc.particle_shift()
        errs["VPERT"] = assert_close("VPERT", eng.field("VPERT")[fluid], orc.field("VPERT")[fluid], MIXED_VPERT)
    print("mixed per-pass errors:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


@pytest.mark.parametrize("name,case", [(n, _mixed(c)) for n, c in _small_cases() if n != "drop3d"],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_step_parity_mixed(name, case):
    """One whole mixed-precision step (start-up pass + integrate) against
    the FP64 oracle.  The sub-iteration counts may differ: Newmark-Beta
    stops at a residual of 1e-7 relative to its first iterate, which can lie
    below the FP32 noise of the pair sums (the 2 m dam-break column runs to
    the iteration limit in mixed precision)."""
    print(f"[mixed step] {name}: n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    ie, io = eng.integrate_step(), orc.integrate_step()
    print(f"sub-iterations: engine {ie['iterations']}, oracle {io['iterations']}", flush=True)
    rho0 = case.settings.rho_rest
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("ACC", eng.field("ACC"), orc.field("ACC"))):
        errs[f] = assert_close(f, g, r, MIXED_STEP)
    if case.settings.ale:
        # the step's pre-passes run Detect_Surface loops 2-3 and
        # particle_shift as one fused kernel; the pass test above covers the
        # separate ones
        fluid = slice(case.n_bound, case.n)
        assert np.array_equal(eng.field("SURFZONE"), orc.field("SURFZONE"))
        for f in ("CURVE", "WOCCL"):
            errs[f] = assert_close(f, eng.field(f)[fluid], orc.field(f)[fluid], MIXED_STEP)
        errs["VPERT"] = assert_close("VPERT", eng.field("VPERT")[fluid], orc.field("VPERT")[fluid], MIXED_VPERT)
    print("mixed per-step errors:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


def test_step_parity_mixed_droplet():
    case = _perturb(cases.droplet_3d(diameter=1.0e-3, dx=5e-5, precision=abi.PREC_MIXED), seed=4)
    print(f"[mixed step] droplet n={case.n}", flush=True)
    eng = _engine(case)
    orc = _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    eng.integrate_step()
    orc.integrate_step()
    rho0 = case.settings.rho_rest
    pairs = {
        "X-x0": (eng.field("X") - case.x, orc.field("X") - case.x),
        "V": (eng.field("V"), orc.field("V")),
        "RHO-rho0": (eng.field("RHO") - rho0, orc.field("RHO") - rho0),
        "P": (eng.field("P"), orc.field("P")),
        "ACC": (eng.field("ACC"), orc.field("ACC")),
        "RRHO": (eng.field("RRHO"), orc.field("RRHO")),
    }
    errs = {k: assert_close(k, g, r, MIXED_STEP) for k, (g, r) in pairs.items()}
    print("mixed per-step errors:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


def test_step_parity_mixed_ale_2d():
    """A 2D droplet with ALE shifting in mixed precision: two steps, whose
    pre-passes run Detect_Surface loops 2-3 and particle_shift as one fused
    2D kernel, against the FP64 oracle."""
    case = _perturb(cases.droplet_2d(diameter=2e-3, dx=5e-5, precision=abi.PREC_MIXED), seed=9, dv=0.01, drho=1e-4)
    case.settings.ale = 1
    case.settings.cfl = 0.3  # Newmark-Beta converges (10-11 sub-iterations) instead of running to the limit
    print(f"[mixed step, 2D ALE] droplet n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(2):
        ie, io = eng.integrate_step(), orc.integrate_step()
        print(f"sub-iterations: engine {ie['iterations']}, oracle {io['iterations']}", flush=True)
    rho0 = case.settings.rho_rest
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("ACC", eng.field("ACC"), orc.field("ACC")),
                    ("CURVE", eng.field("CURVE"), orc.field("CURVE")),
                    ("WOCCL", eng.field("WOCCL"), orc.field("WOCCL"))):
        errs[f] = assert_close(f, g, r, MIXED_STEP)
    assert np.array_equal(eng.field("SURFZONE"), orc.field("SURFZONE"))
    errs["VPERT"] = assert_close("VPERT", eng.field("VPERT"), orc.field("VPERT"), MIXED_VPERT)
    print("mixed 2D ALE errors:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


def test_steps_parity_mixed_jet3d_ale():
    """The 3D jet column as the throughput runs it (mixed precision, ALE
    shifting, fused surface/shift pre-pass, surface tension), unperturbed,
    over six steps (start-up pass + 5) against the FP64 oracle: the per-step
    FP32 pair noise must not grow over the steps beyond MIXED_STEP."""
    case = cases.jet_column_3d(n_target=6000, diameter=0.6e-3, dx=5e-5, precision=abi.PREC_MIXED)
    assert case.settings.ale
    print(f"[mixed steps, 3D ALE] jet n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for k in range(5):
        ie, io = eng.integrate_step(), orc.integrate_step()
        print(f"step {k}: dt engine {ie['dt']:.6e} oracle {io['dt']:.6e}, sub-iterations: engine "
              f"{ie['iterations']}, oracle {io['iterations']}", flush=True)
        assert ie["dt"] == pytest.approx(io["dt"], rel=MIXED_STEP)
    rho0 = case.settings.rho_rest
    fluid = slice(case.n_bound, case.n)
    errs = {}
    for f, g, r in (("X-x0", eng.field("X") - case.x, orc.field("X") - case.x),
                    ("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0),
                    ("ACC", eng.field("ACC"), orc.field("ACC")),
                    ("CURVE", eng.field("CURVE")[fluid], orc.field("CURVE")[fluid]),
                    ("WOCCL", eng.field("WOCCL")[fluid], orc.field("WOCCL")[fluid])):
        errs[f] = assert_close(f, g, r, MIXED_STEP)
    assert np.array_equal(eng.field("SURFZONE"), orc.field("SURFZONE"))
    errs["VPERT"] = assert_close("VPERT", eng.field("VPERT")[fluid], orc.field("VPERT")[fluid], MIXED_VPERT)
    print("mixed 3D ALE errors after 5 steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)


def test_neighbours_full_droplet():
    """Bit-exact sorted neighbour sets on the 268k-particle droplet IC."""
    case = cases.droplet_3d()
    print(f"[neighbours] droplet n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.build_neighbours()
    orc.build_neighbours()
    off_e, idx_e = eng.neighbours()
    off_o, idx_o, _ = orc.neighbours()
    assert same_neighbour_sets(off_e, idx_e, off_o, idx_o)


@pytest.mark.parametrize("dim,H_fac,targets", [(2, 1.3, 256), (2, 2.0, 64), (3, 1.3, 256), (3, 2.0, 1024)])
def test_neighbours_irregular(dim, H_fac, targets):
    """Bit-exact neighbour sets on irregular clouds, at several tile-block
    sizes (the wave's reach box, the two FP32 masks and the target's own
    slot depend on how targets and candidates fall into blocks and tiles)."""
    case = cases.irregular_cloud(dim, H_fac=H_fac, seed=dim * 10 + int(H_fac * 10))
    print(f"[neighbours irregular] dim={dim} H_fac={H_fac} targets={targets} n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.set_block_targets(targets)
    eng.build_neighbours()
    orc.build_neighbours()
    off_e, idx_e = eng.neighbours()
    off_o, idx_o, _ = orc.neighbours()
    assert same_neighbour_sets(off_e, idx_e, off_o, idx_o)


def test_dam_break_drift_fp64():
    """2D dam break (dx = 0.04): 40 steps, FP64 engine vs FP64 oracle."""
    case = cases.dam_break_2d(dx=0.04)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for _ in range(40):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"]
    for f in ("X", "V", "RHO", "P"):
        e = assert_close(f, eng.field(f), orc.field(f), 1e-7)
        print(f"drift {f}: {e:.2e}", flush=True)


def _inlet_cases():
    out = [("fixed", cases.pipe_inlet_3d(mode=0)), ("dynamic", cases.pipe_inlet_3d(mode=1))]
    for mode, name in ((0, "fixed_capped"), (1, "dynamic_capped")):
        capped = cases.pipe_inlet_3d(mode=mode)
        ncol = capped.inlets[0][1].shape[0]
        # the insertion comes after one layer is deleted; max_points stops
        # it part way through the columns
        capped.max_points = capped.n - ncol // 2
        out.append((name, capped))
    for mode, name in ((0, "fixed_rk4"), (1, "dynamic_rk4")):
        rk = cases.pipe_inlet_3d(mode=mode)
        rk.settings.integrator = abi.RUNGE_KUTTA
        out.append((name, rk))
    return out


@pytest.mark.parametrize("name,case", _inlet_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_inlet_parity_fp64(name, case):
    """Pipe inlet (src/shapes/inlet.cpp: buffer, column advance and
    insertion), the pipe exit (Check_Pipe_Outlet) and the deletion plane
    (src/Integration.cpp update_data) over 6 steps that see one deletion
    and one insertion: counts, block ranges and particle types exactly,
    the fields to the FP64 step tolerance, in host order."""
    print(f"[inlet parity] {name}: n={case.n} max_points={case.max_points}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    ns = []
    for step in range(6):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], f"step {step}: {ie} vs {io}"
        assert abs(ie["dt"] - io["dt"]) <= 1e-10 * io["dt"]
        assert eng.n == orc.n, f"step {step}: {eng.n} vs {orc.n} particles"
        assert tuple(eng.block_range(0)) == tuple(orc.block_range(0))
        assert np.array_equal(eng.field("B"), orc.field("B")), f"step {step}: particle types"
        assert np.array_equal(eng.field("PID"), orc.field("PID")), f"step {step}: particle ids"
        (eb, ef), (ob, of) = eng.inlet(0), orc.inlet(0)
        assert np.array_equal(eb, ob) and np.array_equal(ef, of), f"step {step}: inlet columns"
        ns.append(eng.n)
        for f in ("X", "V", "RHO", "P"):
            assert_close(f, eng.field(f), orc.field(f), FP64_STEP)
    print("particle counts:", ns, flush=True)
    assert min(ns) < case.n, "no deletion seen"
    assert max(ns) > min(ns), "no insertion seen"


def test_inlet_parity_mixed():
    """The pipe inlet under the mixed-precision pair passes: the integer
    decisions (insertion, deletion, types) exactly, fields to the mixed
    tolerance."""
    case = cases.pipe_inlet_3d(mode=0, precision=abi.PREC_MIXED)
    print(f"[inlet parity mixed] n={case.n}", flush=True)
    eng, orc = _engine(case), _oracle(case)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for step in range(6):
        eng.integrate_step()
        orc.integrate_step()
        assert eng.n == orc.n
        assert np.array_equal(eng.field("B"), orc.field("B")), f"step {step}: particle types"
    rho0 = case.settings.rho_rest
    for f, g, r in (("V", eng.field("V"), orc.field("V")),
                    ("RHO-rho0", eng.field("RHO") - rho0, orc.field("RHO") - rho0)):
        print(f"{f}: {assert_close(f, g, r, 1e-3):.2e}", flush=True)
