"""The IPT restatement (oracle/ipt_oracle.hpp) against what the mathematics
fixes, on the CPU:

- a particle moving with a uniform gas (no gravity) feels no drag: it
  crosses a structured mesh in a straight line at the gas speed, one step
  per cell, and its record (the state Terminate_Particle keeps, before the
  last time increment) is closed form;
- in still gas under gravity it reaches the terminal velocity where the
  drag of GetCd balances gravity (found here by root finding);
- the reference's ray-triangle test (MollerTrumbore with the edge v2 - v1)
  reports a crossing exactly in the region (1-u, u-v, v) of its
  parameters, computed from the barycentric coordinates, and its distance
  is the ray-plane distance;
- in 3D, tracks in a uniform gas that the test lets through end on the far
  boundary, undeflected."""
import numpy as np
import pytest
from scipy.optimize import brentq

from fjsph_amd import abi, cases


@pytest.fixture(scope="module")
def O():
    from oracle.pyoracle import Oracle
    return Oracle


def _settings(dim, dx, grav):
    s = abi.default_settings(dim)
    s.particle_step = dx
    s.aero_source = abi.MESH_INFL
    for k in range(3):
        s.grav[k] = grav[k] if k < len(grav) else 0.0
    return s


@pytest.mark.parametrize("order", [1, 2])
def test_uniform_flow_crossing(O, order):
    h, nx, ny, U = 0.01, 20, 10, 3.0
    o = O(_settings(2, 1e-4, (0.0, 0.0)), nthreads=2)
    o.set_mesh(cases.quad_mesh(nx, ny, h, [U, 0.0]))
    o.set_ipt(abi.default_ipt(using_ipt=1, eq_order=order))
    rng = np.random.default_rng(3)
    x = np.stack([rng.uniform(0.0, 0.19, 30), rng.uniform(0.001, 0.099, 30)], axis=1)
    i0 = np.floor(x[:, 0] / h).astype(int)
    cell = np.floor(x[:, 1] / h).astype(int) * nx + i0
    t0 = 0.25
    r = o.ipt_track(x, np.tile([U, 0.0], (30, 1)), cell, t0=t0)
    assert np.all(r["status"] == abi.IPT_SUCCESS) and np.all(r["cell"] == -2)
    assert np.array_equal(r["steps"], nx - i0)
    assert np.allclose(r["x"][:, 0], nx * h, rtol=0, atol=1e-12)
    assert np.array_equal(r["x"][:, 1], x[:, 1])
    # the second-order update reduces to v algebraically, not bit for bit
    assert np.allclose(r["v"][:, :2], np.tile([U, 0.0], (30, 1)), rtol=0, atol=1e-12)
    # the time of the last crossing is not yet added when the track ends
    assert np.allclose(r["t"], t0 + (nx * h - x[:, 0] - h) / U, rtol=0, atol=1e-12)


@pytest.mark.parametrize("order", [1, 2])
def test_terminal_velocity(O, order):
    s = _settings(2, 1e-5, (0.0, -9.81))
    o = O(s, nthreads=1)
    o.set_mesh(cases.quad_mesh(2, 1000, 0.005, [0.0, 0.0], rho=1.2))
    o.set_ipt(abi.default_ipt(using_ipt=1, eq_order=order))
    r = o.ipt_track(np.array([[0.0051, 4.99]]), np.array([[0.0, -0.1]]), np.array([999 * 2 + 1]))
    assert r["status"][0] == abi.IPT_SUCCESS and r["steps"][0] == 1000
    d = (6.0 * s.rho_rest * s.particle_step ** 2 / (np.pi * s.rho_rest)) ** (1.0 / 3.0)  # Set_Values ipt_diam
    cd = lambda Re: (1.0 + 0.197 * Re ** 0.63 + 2.6e-04 * Re ** 1.38) * (24.0 / (Re + 0.00001))
    vt = brentq(lambda v: v * v * 3 * cd(2 * 1.2 * v * d / s.mu_g) * 1.2 / (4 * d * s.rho_rest) - 9.81, 1e-6, 100.0)
    assert r["v"][0][0] == 0.0 and r["x"][0][0] == 0.0051
    assert abs(-r["v"][0][1] - vt) < 1e-4 * vt


def test_moller_region_and_distance(O):
    rng = np.random.default_rng(8)
    hits = misses = 0
    for _ in range(400):
        v0, v1, v2 = rng.uniform(-1, 1, (3, 3))
        o = rng.uniform(-1, 1, 3) + np.array([0, 0, 3.0])
        # aim at a point of the triangle's plane, inside or outside it
        lam = rng.uniform(-0.4, 1.2, 3)
        lam /= lam.sum()
        target = lam[0] * v0 + lam[1] * v1 + lam[2] * v2
        d = (target - o) * rng.uniform(0.5, 2.0)
        hit, dt, denom = O.moller(v0, v1, v2, o, d)
        # the reference's parameters: P = v0 + u (v1 - v0) + v (v2 - v1)
        u, v = lam[1] + lam[2], lam[2]
        inside = (u >= 0) and (u <= 1) and (v >= 0) and (u + v <= 1)
        if min(abs(u), abs(u - 1), abs(v), abs(u + v - 1)) < 1e-9:
            continue
        assert hit == inside
        if hit:
            hits += 1
            n = np.cross(v1 - v0, v2 - v0)
            assert dt == pytest.approx(np.dot(v0 - o, n) / np.dot(d, n), rel=1e-9)
            assert denom == (1.0 if dt > np.finfo(float).eps else -1.0)
        else:
            misses += 1
    assert hits > 50 and misses > 50


def test_uniform_flow_3d_tracks_end_on_the_far_boundary(O):
    lo, hi = np.zeros(3), np.array([0.2, 0.1, 0.1])
    m = cases.kuhn_mesh(lo, hi, 6000, speed=3.0, rho_g=1.2, p_inf=1e5)
    m.vert_vel = np.tile([3.0, 0.0, 0.0], (m.nverts, 1))
    o = O(_settings(3, 1e-4, (0.0, 0.0, 0.0)), nthreads=4)
    o.set_mesh(m)
    o.set_ipt(abi.default_ipt(using_ipt=1))
    from tests.test_oracle import _kuhn_cell
    n = np.array([len(np.unique(m.verts[:, k])) - 1 for k in range(3)])
    rng = np.random.default_rng(1)
    x = rng.uniform([0.01, 0.01, 0.01], [0.05, 0.09, 0.09], size=(40, 3))
    r = o.ipt_track(x, np.tile([3.0, 0.0, 0.0], (40, 1)), _kuhn_cell(lo, (hi - lo) / n, n, x))
    ok = r["status"] == abi.IPT_SUCCESS
    assert ok.sum() >= 10
    assert np.allclose(r["x"][ok, 0], 0.2, rtol=0, atol=1e-12)
    assert np.array_equal(r["x"][ok, 1:], x[ok, 1:])
    assert np.allclose(r["v"][ok], [3.0, 0.0, 0.0], rtol=0, atol=1e-12)
    # the others stop at once: the reference's ray-triangle test misses their exit face
    assert np.all(r["steps"][~ok] == 1) and np.all(r["status"][~ok] == abi.IPT_FAILED)
