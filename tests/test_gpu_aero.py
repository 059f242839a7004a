"""Induced-pressure and skin-friction aerodynamics (src/Aero.h:106-263) on the
GPU against the FP64 oracle, through the C-ABI.

The per-particle force is evaluated in FP64 in the Newmark tail
(k_nb_update) in both precision modes; its inputs (normal, scaled curvature,
lambda, the gas cell values) come from the pre-passes.  Tolerances as in
test_gpu_parity: FP64 1e-9 per pass and 1e-8 over steps, mixed 1e-5 per
step, iteration counts equal.  The oracle's two models are pinned by the
flat-sheet closed forms of tests/test_oracle_pins.py.
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close

pytestmark = pytest.mark.gpu


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


def _drop(dim, aero, precision=abi.PREC_FP64, seed=21):
    if dim == 3:
        c = cases.droplet_3d(diameter=1.0e-3, dx=5e-5, precision=precision)
    else:
        c = cases.droplet_2d(diameter=2e-3, dx=1e-4, precision=precision)
    c.settings.aero_case = aero
    # a gas velocity with a cross component, so the skin friction and the
    # pressure coefficient see every angle around the drop
    c.settings.v_inf[1] = 6.0
    return _perturb(c, seed)


AEROS = [(abi.INDUCED_PRESSURE, "induced"), (abi.SKIN_FRIC, "skin")]


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("aero,label", AEROS, ids=[a[1] for a in AEROS])
def test_aero_pass_and_steps_fp64(dim, aero, label):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    case = _drop(dim, aero)
    print(f"[aero {label}] dim={dim} n={case.n}", flush=True)
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    for o in (eng, orc):
        o.build_neighbours()
    npd = orc.dsph_prestep()
    eng.dsph_prestep()
    for o in (eng, orc):
        o.aero_velocity()
        o.detect_surface()
        o.dissipation_terms()
        o.acc_rrho(npd)
    af_o = orc.field("AF")
    assert np.abs(af_o).max() > 0 and (np.abs(af_o).sum(axis=1) > 0).sum() > 20
    for f in ("AF", "ACC", "RRHO"):
        assert_close(f, eng.field(f), orc.field(f), 1e-9)
    eng.close()
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for step in range(3):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], (step, ie, io)
    for f in ("X", "V", "RHO", "ACC", "AF"):
        assert_close(f, eng.field(f), orc.field(f), 1e-8)
    eng.close()


@pytest.mark.parametrize("aero,label", AEROS, ids=[a[1] for a in AEROS])
def test_aero_steps_mixed(aero, label):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    case = _drop(3, aero, precision=abi.PREC_MIXED, seed=22)
    print(f"[aero {label} mixed] n={case.n}", flush=True)
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for step in range(2):
        ie, io = eng.integrate_step(), orc.integrate_step()
        assert ie["iterations"] == io["iterations"], (step, ie, io)
    for f in ("X", "V", "RHO", "ACC"):
        assert_close(f, eng.field(f), orc.field(f), 1e-5)
    eng.close()
