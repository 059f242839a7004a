"""Host-side checks of the engine library that need no GPU: the C-ABI loads
and exports every symbol include/fjsph_engine.h declares, the ctypes struct
layouts match the C ones, settings defaults / Set_Values / the para-file
reader behave like the reference's, and entry points fail cleanly."""
import ctypes as C
import os
import re
import subprocess
import textwrap

import numpy as np
import pytest

from fjsph_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fjsph_engine.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fjsph_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = engine.load()
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(names) == set(engine.SIGNATURES), "ctypes signature table out of sync with the header"


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(textwrap.dedent(f"""
        #include <stdio.h>
        #include <stddef.h>
        #include "{HEADER}"
        int main(void) {{
            printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(fjsph_settings), sizeof(fjsph_derived),
                   sizeof(fjsph_block), sizeof(fjsph_step_info), offsetof(fjsph_settings, dim),
                   offsetof(fjsph_step_info, iterations));
            return 0;
        }}
    """))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = [int(v) for v in out]
    assert got == [C.sizeof(abi.Settings), C.sizeof(abi.Derived), C.sizeof(abi.Block), C.sizeof(abi.StepInfo),
                   abi.Settings.dim.offset, abi.StepInfo.iterations.offset]


@pytest.mark.parametrize("dim", [2, 3])
def test_defaults_agree(dim):
    c_side = engine.settings_default(dim)
    py_side = abi.default_settings(dim)
    assert c_side.as_dict() == py_side.as_dict()


@pytest.mark.parametrize("dim,hfac", [(2, 1.3), (3, 2.0), (3, 1.3)])
def test_derive_matches_oracle(oracle_mod, dim, hfac):
    s = abi.default_settings(dim)
    s.particle_step = 4e-5
    s.H_fac = hfac
    s.sig = 0.0728
    d_eng = engine.derive(s)
    d_orc = oracle_mod.derive(s)
    for name, _ in abi.Derived._fields_:
        a, b = getattr(d_eng, name), getattr(d_orc, name)
        assert a == pytest.approx(b, rel=1e-13, abs=0.0), name


def test_derive_rejects_bad_settings():
    s = abi.default_settings(3)
    with pytest.raises(engine.EngineError):
        engine.derive(s)  # particle_step unset (-1), as GetInput rejects (src/IO.cpp:604-608)
    s.particle_step = 1e-3
    s.dim = 4
    with pytest.raises(engine.EngineError):
        engine.derive(s)


PARA = """\
                          Reference dispersed density: 998.2
                        Reference dispersed viscosity: 1.0e-3   # a comment
                            Reference surface tension: 0.0728
# SPH parameters ------------------------------------: -
                                 SPH maximum timestep: 2.5e-4
                                    SPH CFL condition: 0.3
                                  SPH initial spacing: 0.02
                          SPH smoothing length factor: 1.3
                                 SPH aerodynamic case: Gissler
                                   SPH gravity vector: 0.0,-9.81
                              SPH freestream velocity: 20,1.5
                                   SPH speed of sound: 62.64
                               SPH integration solver: Runge-Kutta
                      SPH Newmark-Beta iteration limit: 15
                                    Unknown key ignored: 5
"""


def test_para_parser(tmp_path):
    p = tmp_path / "para"
    p.write_text(PARA)
    s = engine.parse_para(str(p), 2)
    assert s.rho_rest == 998.2 and s.mu == 1.0e-3 and s.sig == 0.0728
    assert s.delta_t_max == 2.5e-4 and s.cfl == 0.3 and s.particle_step == 0.02 and s.H_fac == 1.3
    assert s.aero_case == abi.GISSLER and s.integrator == abi.RUNGE_KUTTA and s.max_subits == 15
    assert list(s.grav)[:2] == [0.0, -9.81] and list(s.v_inf)[:2] == [20.0, 1.5]
    assert s.speed_sound == 62.64
    bad = tmp_path / "bad"
    bad.write_text("SPH aerodynamic case: Warp_drive\n")
    with pytest.raises(engine.EngineError):
        engine.parse_para(str(bad), 2)


def test_create_fails_cleanly_without_device():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a HIP device is visible")
    s = abi.default_settings(3)
    s.particle_step = 1e-3
    with pytest.raises(engine.EngineError, match="status 6"):
        engine.Engine(s)


def test_create_rejects_unbuilt_options():
    lib = engine.load()
    s = abi.default_settings(3)
    s.particle_step = 1e-3
    s.aero_source = abi.VLM_INFL
    h = C.c_void_p()
    assert lib.fjsph_create(C.byref(s), 0, C.byref(h)) in (abi.ERR_UNSUPPORTED, abi.ERR_NO_DEVICE)
    s.aero_source = abi.CONST_VEL
    s.precision = 7
    assert lib.fjsph_create(C.byref(s), 0, C.byref(h)) == abi.ERR_ARG
    assert lib.fjsph_create(None, 0, C.byref(h)) == abi.ERR_ARG


def test_cases_are_well_formed():
    from fjsph_amd import cases

    c = cases.dam_break_2d()
    assert c.n - c.n_bound == 5000 and 1400 < c.n_bound < 1700
    assert np.all(c.b[:c.n_bound] == abi.BOUND) and np.all(c.b[c.n_bound:] == abi.FREE)
    d = cases.droplet_3d()
    assert 260_000 < d.n < 275_000
    j = cases.jet_column_3d(n_target=100_000)
    assert abs(j.n - 100_000) < 2_000
    assert np.allclose(j.v[:, 0], 10.0)
