"""FindCell's particles in no cell on the GPU (the boundary test of
k_lost_classify, the deletion and the rebuilt neighbour lists) against the
oracle, step by step in FP64: droplets straddling the mesh boundary

- above a far-field top: the surface particles in no cell are deleted;
- above a wall top: they keep going with the internal flag;
- right of a 2D mesh, where the one-sided edge test sees no crossing: the
  failure count rises by one per lookup (two per step), and the particles
  are deleted at the lookup after it passes 10;
- on a mesh of more than 4096 cells (the radius search for the 500 nearest
  centres), and in 3D (Crossings3D against tetrahedral faces).

Particle counts, ids, failure counts, internal flags, cells and the
counters agree exactly; positions, velocities and densities to 1e-10."""
import numpy as np
import pytest

from fjsph_amd import abi, cases
from tests.parity import assert_close
from tests.test_lost import _mark

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _droplet_2d(nx, ny, h, centre, wall_top=False):
    case = cases.droplet_2d(diameter=4e-3, dx=2e-4, centre=centre, speed=3.0)
    case.v[:] = [3.0, 0.0]
    case.settings.aero_source = abi.MESH_INFL
    m = cases.quad_mesh(nx, ny, h, [3.0, 0.0], p=case.settings.p_ref, rho=1.2)
    if wall_top:
        m = _mark(m, lambda c: c[:, 1] > ny * h - 1e-12, -1)
    case.mesh = m
    return case


def _droplet_3d():
    case = cases.droplet_3d(diameter=0.8e-3, dx=5e-5, precision=abi.PREC_FP64, speed=3.0)
    case.settings.aero_source = abi.MESH_INFL
    lo, hi = case.x.min(axis=0), case.x.max(axis=0)
    c = 0.5 * (lo + hi)
    mlo = lo - 0.6e-3
    mhi = np.array([hi[0] + 0.6e-3, hi[1] + 0.6e-3, c[2] + 0.1e-3])
    case.mesh = cases.kuhn_mesh(mlo, mhi, 20000, speed=3.0, rho_g=1.2, p_inf=case.settings.p_ref)
    case.mesh.vert_vel = np.tile([3.0, 0.0, 0.0], (case.mesh.nverts, 1))
    return case


CASES = {
    "far_top": lambda: _droplet_2d(10, 5, 0.01, (0.05, 0.05)),
    "wall_top": lambda: _droplet_2d(10, 5, 0.01, (0.05, 0.05), wall_top=True),
    "right": lambda: _droplet_2d(5, 10, 0.01, (0.05, 0.05)),
    "far_top_5000_cells": lambda: _droplet_2d(100, 50, 0.002, (0.1, 0.1)),
    "far_top_3d": _droplet_3d,
}


@pytest.mark.parametrize("name,steps", [("far_top", 4), ("wall_top", 4), ("right", 8),
                                        ("far_top_5000_cells", 3), ("far_top_3d", 2)])
def test_lost_particles_match_the_oracle(name, steps):
    from fjsph_amd.engine import Engine
    from oracle.pyoracle import Oracle
    case = CASES[name]()
    print(f"[lost] {name}: n={case.n}, mesh cells {case.mesh.ncells}, {steps} steps", flush=True)
    eng, orc = Engine(case), Oracle(case, nthreads=8)
    eng.integrate_no_update()
    orc.integrate_no_update()
    for step in range(steps):
        eng.integrate_step()
        orc.integrate_step()
        ie, io = eng.lost_info(), orc.lost_info()
        print(f"  step {step}: n {eng.n} / {orc.n}, engine {ie}, oracle {io}", flush=True)
        assert eng.n == orc.n, f"step {step}: particle count"
        assert ie == io, f"step {step}: lookup counters"
        for f in ("PID", "NFAILED", "INTERNAL", "CELLID", "B"):
            assert np.array_equal(eng.field(f), orc.field(f)), f"step {step}: {f}"
        for f in ("X", "V", "RHO"):
            assert_close(f"step {step}: {f}", eng.field(f), orc.field(f), TOL)
    info = eng.lost_info()
    if name == "wall_top":
        assert info["deleted"] == 0 and info["internal"] > 0 and eng.field("INTERNAL").sum() > 0
    elif name == "right":
        assert info["deleted"] > 0 and info["failed"] >= 11 * info["deleted"]
    else:
        assert info["deleted"] > 0 and eng.n < case.n
