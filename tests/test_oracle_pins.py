"""Pinning tests of the oracle's pair terms, boundary solvers and rollback
(CPU only).

Each term is held to a result fixed by the mathematics, not by its own
formula: the continuum limit of an SPH sum on a regular interior lattice
(where the lattice sums are accurate to well under 1% at H = 2 dx), a
closed-form geometry, or an exact weighted mean.  A dropped term, a wrong
sign or a wrong factor fails one of them.

  term (reference file:line)                          pinned by
  --------------------------------------------------  --------------------------------------------------
  laminar viscosity (src/Kernel.h:262-269,            Morris' estimate tends to nu lap(v): on v_x = a y^2
    src/Resid.cpp:311-313)                            the interior acceleration is 2 a nu (1%)
  artificial viscosity (src/Kernel.h:217-244,         Monaghan's viscosity tends to nu_AV (lap v +
    src/Shifting.cpp:126-186)                         2 grad div v), nu_AV = alpha c H / (2 (D + 2)),
                                                      for a flow whose pairs all approach (6%); no
                                                      force for separating pairs; kinetic energy falls;
                                                      total momentum kept
  ALE momentum and continuity (src/Kernel.h:179-195,  (du . grad) v = A du for v = A x and a uniform
    src/Resid.cpp:325-341, 446-455)                   shift du (2%); drho/dt = -rho div v + du . grad rho
                                                      (0.5%); du = 0 equals the plain continuity exactly
  curvature (src/Geometry.cpp:131-259)                outward normals; kappa R / (D - 1) near 1 on discs
                                                      and spheres, approaching 1 as R / dx grows, and
                                                      kappa ~ 1 / R
  Gissler occlusion (src/Geometry.cpp:236-273)        a two-layer sheet across the stream: the upstream
                                                      layer exposed (0), the downstream one shadowed (1);
                                                      at 45 degrees the lattice angles (0, 1/sqrt 2, 1)
  Gissler plate branch (src/Aero.h:37-98)             frac2 = 1: a flat plate of area dx^2 with Cd = 1,
                                                      times the exposed fraction 1 - woccl
  induced pressure (src/Aero.h:106-202)               a two-layer sheet (scaled curvature > 1/4) at frac1 = 1:
                                                      the stagnation pressure 1/2 rho_g U^2 on the exposed
                                                      face (Cp = 1) pushing along the stream, the base
                                                      Cp = 0.075 on the rear face pushing against its
                                                      normal; with a tangential gas velocity W the
                                                      sphere curve Cp = 1 - 9/4 sin^2 theta on the
                                                      normal and the seventh-power skin friction
                                                      1/2 rho_g W^2 0.027 Re_W^(-1/7) on the tangent
  skin friction (src/Aero.h:223-256)                  no force on a face the gas leaves (V . n <= 0);
                                                      on the others the stagnation pressure of the
                                                      normal velocity plus the skin friction of the
                                                      tangential one, blended with the droplet's sphere
                                                      drag by min(1.5 lambda, 1)
  ghost boundary (src/Resid.cpp:122-160)              drho/dt = -rho div v of the surrounding fluid
                                                      (0.2%); near_inlet clamps (src/Newmark_Beta.cpp:
                                                      163-193)
  no-slip (src/Resid.cpp:162-186)                     wall velocity 2 V_wall - U under a uniform fluid
                                                      velocity U (the kernel-weighted mean of a constant)
  DBC (src/Resid.cpp:78-119, src/Newmark_Beta.cpp:    zero boundary acceleration, density held
    145-160)
  block motion table (src/Newmark_Beta.cpp:69-95)     the velocity of the last time strictly before now
  whole runs: Couette and Poiseuille start-up        the analytic profiles the reference ships
    (Examples/Couette, Examples/Poiseuille)           (Examples/*/SteadyState.dat) at t = 5 s and 0.02 s
  Newmark rollback (src/Newmark_Beta.cpp:32-48)       a step started with 8x a stable dt diverges, is
                                                      rolled back once to pn with dt halved, converges,
                                                      and its final positions satisfy the Newmark update
                                                      from pn with the halved dt
"""
import math

import numpy as np
import pytest

from fjsph_amd import abi, cases


@pytest.fixture(scope="module")
def po(oracle_mod):
    return oracle_mod


def _lattice(dim, n_side, H_fac=2.0, mu=0.0, dx=1e-3):
    """A centred square/cube lattice with the pair terms under test isolated:
    no gravity, no artificial viscosity, no delta-SPH, no surface tension."""
    c = cases.fluid_block(dim, n_side, dx=dx, H_fac=H_fac, mu=mu)
    s = c.settings
    s.visc_alpha = 0.0
    s.dsph_delta = 0.0
    s.sig = 0.0
    for k in range(3):
        s.grav[k] = 0.0
    c.x = np.ascontiguousarray(c.x - c.x.mean(axis=0))
    return c


def _interior(c, layers):
    """Particles whose kernel support lies inside the lattice."""
    dx = c.settings.particle_step
    lo, hi = c.x.min(axis=0), c.x.max(axis=0)
    return np.all((c.x > lo + layers * dx - 1e-12) & (c.x < hi - layers * dx + 1e-12), axis=1)


# ------------------------------------------------------- laminar viscosity

@pytest.mark.parametrize("dim,n_side", [(2, 30), (3, 16)])
def test_laminar_viscosity_tends_to_nu_laplacian(po, dim, n_side):
    """Morris (1997): sum_j m_j (mu_i + mu_j) / (rho_i rho_j) (x_ij . grad_i W)
    / (r^2 + eta^2) v_ij -> nu lap(v).  On v_x = a y^2, lap(v) = (2a, 0, 0)."""
    c = _lattice(dim, n_side, mu=1e-3)
    a = 2.0
    c.v = np.zeros_like(c.x)
    c.v[:, 0] = a * c.x[:, 1] ** 2
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    o.acc_rrho(1.0)
    acc = o.field("ACC")
    sel = _interior(c, 5)
    assert sel.sum() > 100
    nu = c.settings.mu / c.settings.rho_rest
    expect = 2.0 * a * nu
    assert np.allclose(acc[sel, 0], expect, rtol=1e-2, atol=0)
    assert np.max(np.abs(acc[sel, 1:])) < 1e-9 * expect


# ---------------------------------------------------- artificial viscosity

def _compressing(c, a=1.0):
    """v = -b x + (a y^2, 0, 0) with b large enough that every neighbour pair
    approaches: (v_j - v_i).(x_j - x_i) = -b r^2 + a (y_i + y_j) dx dy < 0."""
    Y = np.abs(c.x[:, 1]).max()
    b = 2.5 * a * Y
    v = -b * c.x
    v[:, 0] += a * c.x[:, 1] ** 2
    return np.ascontiguousarray(v), b


@pytest.mark.parametrize("dim,n_side", [(2, 30), (3, 16)])
def test_artificial_viscosity_effective_viscosity(po, dim, n_side):
    """Monaghan's viscosity with Pi = -alpha c mu / rho, mu = H v.r / (r^2 +
    eps H^2), tends to nu_AV (lap v + 2 grad div v) with nu_AV = alpha c H /
    (2 (D + 2)) (the continuum limit of the pair sum; Monaghan, Rep. Prog.
    Phys. 68 (2005) 1703).  Here div v is uniform and lap v = (2a, 0, 0)."""
    c = _lattice(dim, n_side)
    s = c.settings
    s.visc_alpha = 0.1
    a = 1.0
    c.v, _ = _compressing(c, a)
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    o.dsph_prestep()
    o.dissipation_terms()
    av = o.field("AVISC")
    d = po.derive(s)
    nu_av = s.visc_alpha * s.speed_sound * d.H / (2 * (dim + 2))
    expect = nu_av * 2.0 * a
    sel = _interior(c, 5)
    assert sel.sum() > 100
    assert np.allclose(av[sel, 0], expect, rtol=0.06, atol=0)
    assert np.max(np.abs(av[sel, 1:])) < 1e-9 * expect
    # a dissipative force: the kinetic energy falls; the pair forces are
    # equal and opposite, so the total momentum is kept
    assert (c.m[:, None] * c.v * av).sum() < 0.0
    total = (c.m[:, None] * av).sum(axis=0)
    assert np.all(np.abs(total) < 1e-12 * (c.m[:, None] * np.abs(av)).sum())


@pytest.mark.parametrize("dim,n_side", [(2, 16), (3, 9)])
def test_artificial_viscosity_off_for_separating_pairs(po, dim, n_side):
    c = _lattice(dim, n_side)
    c.settings.visc_alpha = 0.1
    v, _ = _compressing(c)
    c.v = -v                       # every pair separates
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    o.dsph_prestep()
    o.dissipation_terms()
    assert np.all(o.field("AVISC") == 0.0)


# ------------------------------------------------------------------- ALE

def _ale_lattice(dim, n_side, seed):
    """v = A x, rho = rho0 + b.x with masses m = rho dx^D, so that every
    particle's volume is the lattice cell and the lattice sums are exact
    quadratures; a uniform shift velocity du."""
    c = _lattice(dim, n_side)
    c.settings.ale = 1
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(dim, dim)) * 10.0
    A -= (np.trace(A) / dim - 0.3) * np.eye(dim)   # div v = 0.3 D: comparable to du . grad(rho)
    bvec = rng.normal(size=dim) * 500.0
    du = 0.3 * bvec / np.linalg.norm(bvec) + rng.normal(size=dim) * 0.1
    c.v = np.ascontiguousarray(c.x @ A.T)
    c.rho = 1000.0 + c.x @ bvec
    c.p = np.zeros(c.n)
    c.m = c.rho * c.settings.particle_step ** dim
    return c, A, du, bvec


@pytest.mark.parametrize("dim,n_side", [(2, 24), (3, 14)])
def test_ale_momentum_is_advection_by_the_shift(po, dim, n_side):
    """The ALE momentum term (src/Kernel.h:179-186) estimates
    div(v (x) du) - v div(du) = (du . grad) v: A du for v = A x."""
    c, A, du, _ = _ale_lattice(dim, n_side, seed=5)
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    o.set_field("VPERT", np.tile(du, (c.n, 1)))
    o.acc_rrho(1.0)
    acc = o.field("ACC")
    sel = _interior(c, 5)
    expect = A @ du
    err = np.max(np.abs(acc[sel] - expect)) / np.linalg.norm(expect)
    assert err < 0.02


@pytest.mark.parametrize("dim,n_side", [(2, 24), (3, 14)])
def test_ale_continuity(po, dim, n_side):
    """ALEContinuity and ALECont2ndterm (src/Kernel.h:187-195): drho/dt =
    -rho div(v + du) + div(rho du) = -rho tr(A) + du . grad(rho); with du = 0
    the ALE continuity equals the plain one exactly."""
    c, A, du, bvec = _ale_lattice(dim, n_side, seed=6)
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    o.set_field("VPERT", np.tile(du, (c.n, 1)))
    o.acc_rrho(1.0)
    rr = o.field("RRHO")
    sel = _interior(c, 5)
    expect = -c.rho * np.trace(A) + du @ bvec
    scale = np.max(np.abs(expect[sel]))
    assert abs(du @ bvec) > 0.1 * scale      # the shift term is resolved
    assert np.max(np.abs(rr[sel] - expect[sel])) < 5e-3 * scale
    o.set_field("VPERT", np.zeros((c.n, dim)))
    o.acc_rrho(1.0)
    r_ale = o.field("RRHO")
    c.settings.ale = 0
    o2 = po.Oracle(c, nthreads=4)
    o2.build_neighbours()
    o2.acc_rrho(1.0)
    assert np.array_equal(o2.field("RRHO"), r_ale)


# ------------------------------------------------------------- curvature

def _ball(dim, radius_dx, H_fac=2.0, dx=1e-3):
    s = abi.default_settings(dim)
    s.particle_step = dx
    s.H_fac = H_fac
    s.mu = 0.0
    s.sig = 0.0
    s.visc_alpha = 0.0
    s.aero_case = abi.NO_AERO
    R = radius_dx * dx
    n = int(2 * radius_dx) + 4
    g = (np.arange(n) - (n - 1) / 2) * dx
    mesh = np.meshgrid(*([g] * dim), indexing="ij")
    x = np.stack([m.ravel() for m in reversed(mesh)], axis=1)
    x = x[(x ** 2).sum(axis=1) <= R * R]
    return cases._finish(f"ball_{dim}d", s, np.zeros((0, dim)), x, 0.0, ""), R


@pytest.mark.parametrize("dim,radii", [(2, (10, 20, 40)), (3, (8, 12, 16))])
def test_curvature_of_discs_and_spheres(po, dim, radii):
    """Curvature kappa = sum_j V_j (L (n_j - n_i)) . grad_i W (src/Geometry.cpp:
    215-235) is the divergence of the unit normal: (D - 1) / R on a disc or
    sphere of radius R with outward normals.  Lattice-clipped balls give a
    scattered estimate; its mean over the surface particles lies within 30%
    of (D - 1) / R, tends to it as R / dx grows, and scales as 1 / R."""
    means = []
    for rdx in radii:
        c, R = _ball(dim, rdx)
        o = po.Oracle(c, nthreads=8)
        o.build_neighbours()
        o.dsph_prestep()
        o.detect_surface()
        surf = o.field("SURF") == 1
        nrm = o.field("NORM")[surf]
        rhat = c.x[surf] / np.linalg.norm(c.x[surf], axis=1)[:, None]
        assert np.min((nrm * rhat).sum(axis=1)) > 0.9          # outward unit normals
        k = o.field("CURVE")[surf] * R / (dim - 1)
        assert np.median(k) > 0.0
        means.append(k.mean())
    assert all(0.9 < m < 1.3 for m in means), means
    assert abs(means[-1] - 1.0) < abs(means[0] - 1.0)
    # kappa ~ 1/R: the dimensional means of the smallest and largest ball
    ratio = (means[0] / radii[0]) / (means[-1] / radii[-1])
    assert ratio == pytest.approx(radii[-1] / radii[0], rel=0.25)


# ------------------------------------------------------ Gissler occlusion

def _sheet(dim, tilt, i_interp_fac=0.5):
    """Two lattice layers across a 20 m/s stream along +x (tilted by 45
    degrees towards +y when `tilt`), constant gas velocity."""
    dx = 1e-4
    c = cases.fluid_block(dim, 12, dx=dx, H_fac=1.3, aero=abi.GISSLER, speed=20.0, mu=0.0)
    xs = c.x[c.x[:, 0] < 2 * dx]
    s = c.settings
    s.sig = 0.0
    s.i_interp_fac = i_interp_fac
    for k in range(3):
        s.grav[k] = 0.0
    if tilt:
        s.v_inf[1] = s.v_inf[0]
    c = cases._finish("sheet", s, np.zeros((0, dim)), xs, 0.0, "")
    return c


@pytest.mark.parametrize("dim", [2, 3])
def test_occlusion_of_a_two_layer_sheet(po, dim):
    """woccl = max_j cos(x_i - x_j, V_diff), clamped to [0, 1]
    (src/Geometry.cpp:236-273): the fraction of a particle's area shadowed
    by an upstream neighbour."""
    for tilt in (False, True):
        c = _sheet(dim, tilt)
        o = po.Oracle(c, nthreads=4)
        o.build_neighbours()
        o.dsph_prestep()
        o.aero_velocity()
        o.detect_surface()
        w = o.field("WOCCL")
        assert np.all(o.field("LAM_NB") < c.settings.lam_cutoff)
        assert np.all((w >= 0.0) & (w <= 1.0))
        dx = c.settings.particle_step
        front = c.x[:, 0] < dx
        if not tilt:
            assert np.all(w[front] == 0.0)
            assert np.allclose(w[~front], 1.0, rtol=0, atol=1e-15)
        else:
            # with the stream along (1, 1, 0): a particle with a lattice
            # neighbour at -y is shadowed at 45 degrees (front layer), one
            # with a neighbour at (-x, -y) straight on (back layer); the
            # lowest row lacks those neighbours
            ylo = c.x[:, 1] < c.x[:, 1].min() + 0.5 * dx
            r2 = 1.0 / math.sqrt(2.0)
            assert np.all(w[front & ylo] == 0.0)
            assert np.allclose(w[front & ~ylo], r2, rtol=0, atol=1e-12)
            assert np.allclose(w[~front & ylo], r2, rtol=0, atol=1e-12)
            assert np.allclose(w[~front & ~ylo], 1.0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_gissler_plate_branch(po, dim):
    """With frac2 = min(lambda_nb interp_fac, 1) = 1 the Gissler force is the
    drag of a flat plate, Cd = 1, of area dx^D-1 (the particle's lattice
    face), times the unshadowed fraction 1 - woccl: 1/2 rho_g U^2 dx^(D-1) / m
    on the exposed layer, nothing on the shadowed one."""
    c = _sheet(dim, False, i_interp_fac=0.1)
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    npd = o.dsph_prestep()
    o.aero_velocity()
    o.detect_surface()
    o.dissipation_terms()
    o.acc_rrho(npd)
    lam = o.field("LAM_NB")
    assert np.all(lam * 10.0 >= 1.0)
    af = o.field("AF")
    dx = c.settings.particle_step
    front = c.x[:, 0] < dx
    rho_g = 1.2
    U = 20.0
    expect = 0.5 * rho_g * U * U * dx ** (dim - 1) / c.m[front]
    assert np.allclose(af[front, 0], expect, rtol=1e-12, atol=0)
    assert np.all(af[front, 1:] == 0.0)
    assert np.all(af[~front] == 0.0)


# ---------------------------------------------------------- boundary solvers

def _embedded_wall(dim, n_side, solver, seed=3, no_slip=0, vel=(0.0, 0.0, 0.0)):
    """A lattice whose central particle (or a slab through the centre) is a
    boundary block, inside fluid moving with v = A x."""
    c = _lattice(dim, n_side)
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(dim, dim)) * 10.0
    centre = np.argmin(np.linalg.norm(c.x, axis=1))
    order = np.r_[centre, np.delete(np.arange(c.n), centre)]
    x = c.x[order]
    v = x @ A.T
    n = c.n
    b = np.full(n, abi.FREE, np.int32)
    b[0] = abi.BOUND
    rho = np.full(n, 1000.0)
    m = rho * c.settings.particle_step ** dim
    blocks = [dict(first=0, last=1, is_fluid=0, bound_solver=solver, no_slip=no_slip, vel=vel),
              dict(first=1, last=n, is_fluid=1)]
    return c.settings, x, v, rho, m, b, blocks, A


@pytest.mark.parametrize("dim,n_side", [(2, 15), (3, 11)])
def test_ghost_boundary_continuity(po, dim, n_side):
    """Boundary_Ghost: drho/dt = -rho sum_j V_j v_ji . grad W -> -rho div v."""
    s, x, v, rho, m, b, blocks, A = _embedded_wall(dim, n_side, abi.GHOST)
    o = po.Oracle(s, nthreads=2)
    o.upload_state(x, v, rho, np.zeros(len(rho)), m, b, 1)
    o.set_blocks(blocks)
    o.build_neighbours()
    o.boundary_update()
    rr = o.field("RRHO")[0]
    expect = -1000.0 * np.trace(A)
    assert rr == pytest.approx(expect, rel=2e-3)


def test_ghost_near_inlet_clamp(po):
    """A ghost particle that sees only BUFFER (and boundary) particles is
    'near the inlet': its density update is clamped to at least rho_rest and
    its rate to at least 0 (src/Newmark_Beta.cpp:163-178), while the same
    particle next to FREE fluid is not."""
    s, x, v, rho, m, b, blocks, A = _embedded_wall(2, 9, abi.GHOST)
    # an expanding fluid: drho/dt < 0 at the wall particle
    v = np.ascontiguousarray(0.5 * x / np.abs(x).max())
    results = {}
    for kind in (abi.FREE, abi.BUFFER):
        bb = b.copy()
        bb[1:] = kind
        o = po.Oracle(s, nthreads=2)
        o.upload_state(x, v, rho, np.zeros(len(rho)), m, bb, 1)
        o.set_blocks(blocks)
        o.integrate_no_update()
        results[kind] = (o.field("RHO")[0], o.field("RRHO")[0])
    rho_free, rr_free = results[abi.FREE]
    rho_buf, rr_buf = results[abi.BUFFER]
    assert rr_free < 0.0 and rho_free < 1000.0
    assert rr_buf == 0.0 and rho_buf == 1000.0


@pytest.mark.parametrize("dim", [2, 3])
def test_no_slip_wall_velocity(po, dim):
    """Set_No_Slip: v_wall <- 2 V_wall - sum_j W v_j / sum_j W over the fluid
    neighbours, which is 2 V_wall - U under a uniform fluid velocity U."""
    c = cases.fluid_block(dim, 8 if dim == 2 else 6, with_walls=True, layers=3, H_fac=1.3, mu=0.0)
    U = np.array([0.3, -0.2, 0.1][:dim])
    V = np.array([0.05, 0.02, -0.01][:dim])
    c.v[c.n_bound:] = U
    c.blocks[0]["no_slip"] = 1
    c.blocks[0]["vel"] = tuple(V) + (0.0,) * (3 - dim)
    o = po.Oracle(c, nthreads=2)
    o.build_neighbours()
    o.boundary_update()
    vb = o.field("V")[:c.n_bound]
    off, idx, _ = o.neighbours()
    near = np.array([np.any(idx[off[k]:off[k + 1]] >= c.n_bound) for k in range(c.n_bound)])
    assert near.sum() > 10 and (~near).sum() > 10
    assert np.allclose(vb[near], 2 * V - U, rtol=0, atol=1e-14)
    assert np.allclose(vb[~near], V, rtol=0, atol=0)


def test_dbc_boundary_is_held(po):
    """Boundary_DBC accumulates its pair terms at the fluid indices
    (RV_[jj.first]) but copies RV_ at the boundary indices, so a DBC wall
    has no acceleration and keeps its density and pressure."""
    c = cases.dam_break_2d(dx=0.1)
    c.blocks[0]["bound_solver"] = abi.DBC
    rng = np.random.default_rng(2)
    c.rho[:c.n_bound] = 1000.0 * (1 + rng.uniform(0, 0.01, c.n_bound))
    o = po.Oracle(c, nthreads=2)
    o.integrate_no_update()
    for _ in range(3):
        o.integrate_step()
    assert np.all(o.field("ACC")[:c.n_bound] == 0.0)
    assert np.array_equal(o.field("RHO")[:c.n_bound], c.rho[:c.n_bound])


def test_block_motion_table(po):
    """The wall velocity is that of the last table time strictly before the
    current time (src/Newmark_Beta.cpp:69-84); before the first time it is 0."""
    c = cases.fluid_block(2, 6, with_walls=True, layers=3, H_fac=1.3, mu=0.0)
    o = po.Oracle(c, nthreads=1)
    times = np.array([0.0, 1.0, 2.0])
    vels = np.array([[0.1, 0, 0], [0.2, 0.05, 0], [0.3, 0, 0]], dtype=np.float64)
    o.lib.orc_set_block_motion(o.h, 0, 3, times.ctypes.data_as(po._D), vels.ctypes.data_as(po._D))
    o.build_neighbours()
    for t, expect in ((0.0, (0.0, 0.0)), (0.5, (0.1, 0.0)), (1.0, (0.1, 0.0)), (1.5, (0.2, 0.05)),
                      (7.0, (0.3, 0.0))):
        o.set_time(t, 1e-4, 1.0)
        o.boundary_update()
        assert np.array_equal(o.field("V")[:c.n_bound], np.tile(expect, (c.n_bound, 1))), t


# -------------------------------------------------------- Newmark rollback

def test_newmark_rollback_halves_dt(po):
    """A step forced to 8x a stable time step (delta_t_min) diverges: after
    max_subits iterations its residual is above the first iterate's, so the
    state rolls back to pn, dt halves and the iterations start again
    (src/Newmark_Beta.cpp:32-48).  The step then converges, and its final
    positions are the Newmark update x = x_n + dt v_n + dt^2 (a_n + a) / 4
    from pn with the halved dt."""
    c = cases.dam_break_2d(dx=0.05)
    o = po.Oracle(c, nthreads=8)
    o.integrate_no_update()
    for _ in range(3):
        o.integrate_step()
    _, dt_stable, _ = o.get_time()
    c.settings.delta_t_min = 8.0 * dt_stable
    o2 = po.Oracle(c, nthreads=8)
    o2.integrate_no_update()
    xn = o2.field("X", abi.STATE_N)
    vn = o2.field("V", abi.STATE_N)
    an = o2.field("ACC", abi.STATE_N)
    info = o2.integrate_step()
    assert info["rollbacks"] == 1
    assert info["dt"] == 0.5 * c.settings.delta_t_min
    assert info["rms_error"] <= c.settings.min_residual
    dt = info["dt"]
    a = o2.field("ACC")
    x = o2.field("X")
    nb = c.n_bound
    expect = xn[nb:] + dt * vn[nb:] + dt * dt * (0.25 * an[nb:] + 0.25 * a[nb:])
    assert np.allclose(x[nb:], expect, rtol=0, atol=1e-13)
    assert np.array_equal(x[:nb], xn[:nb])


# --------------------------------------- induced pressure and skin friction

def _plate_centre(c, dim):
    """Particles of the sheet at least 3 dx from its edges, where the
    normals are along x."""
    dx = c.settings.particle_step
    lo, hi = c.x.min(axis=0), c.x.max(axis=0)
    sel = np.ones(c.n, bool)
    for k in range(1, dim):
        sel &= (c.x[:, k] > lo[k] + 3 * dx) & (c.x[:, k] < hi[k] - 3 * dx)
    return sel


def _run_aero(po, c):
    o = po.Oracle(c, nthreads=4)
    o.build_neighbours()
    npd = o.dsph_prestep()
    o.aero_velocity()
    o.detect_surface()
    o.dissipation_terms()
    o.acc_rrho(npd)
    return o


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("W", [0.0, 12.0])
def test_induced_pressure_flat_sheet(po, dim, W):
    c = _sheet(dim, False, i_interp_fac=0.1)
    s = c.settings
    s.aero_case = abi.INDUCED_PRESSURE
    s.v_inf[1] = W
    o = _run_aero(po, c)
    dx = s.particle_step
    mid = _plate_centre(c, dim)
    nrm = o.field("NORM")
    front = c.x[:, 0] < dx
    assert np.all(o.field("LAM_NB")[mid] * 10.0 >= 1.0)
    assert np.allclose(nrm[mid & front, 0], -1.0, atol=1e-12) and np.allclose(nrm[mid & ~front, 0], 1.0, atol=1e-12)
    # the two faces' opposite normals give a scaled curvature dx kappa near
    # 2, above 0.25: the pressure coefficient is the sphere curve alone
    assert np.all(o.field("NORM_CURVE")[mid] > 0.25)
    af = o.field("AF")
    rho_g, U = 1.2, 20.0
    A, m = dx ** (dim - 1), c.m[0]
    V2 = U * U + W * W
    sin2 = W * W / V2
    # exposed face: Cp = 1 - 9/4 sin^2(theta) along +x, friction along +y
    cp = 1.0 - 2.25 * sin2
    assert np.allclose(af[mid & front, 0], 0.5 * rho_g * V2 * cp * A / m, rtol=1e-9, atol=0)
    mu_g = s.mu_g
    L = dx * (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) if dim == 3 else dx / math.sqrt(math.pi)
    cf = 0.027 / (rho_g * W * L / mu_g + 1e-6) ** (1.0 / 7.0)
    scale = 0.5 * rho_g * V2 * A / m
    assert np.allclose(af[mid & front, 1], 0.5 * rho_g * W * W * cf * A / m, rtol=1e-9, atol=1e-12 * scale)
    # rear face: theta >= 2.4455 for W = 0 (base Cp 0.075 against the normal)
    if W == 0.0:
        assert np.allclose(af[mid & ~front, 0], -0.5 * rho_g * U * U * 0.075 * A / m, rtol=1e-9, atol=0)
        assert np.all(np.abs(af[mid][:, 1:]) <= 1e-9 * np.abs(af[mid, 0]).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_skin_friction_flat_sheet(po, dim):
    c = _sheet(dim, False)
    s = c.settings
    s.aero_case = abi.SKIN_FRIC
    W = 8.0
    s.v_inf[1] = W
    o = _run_aero(po, c)
    dx = s.particle_step
    mid = _plate_centre(c, dim)
    front = c.x[:, 0] < dx
    af = o.field("AF")
    # the gas leaves the upstream face's outward normal (-x): no force
    assert np.all(af[mid & front] == 0.0)
    rear = mid & ~front
    lam = o.field("LAM_NB")[rear]
    frac = np.minimum(1.5 * lam, 1.0)
    rho_g, U = 1.2, 20.0
    A, m = dx ** (dim - 1), c.m[0]
    L = dx * (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) if dim == 3 else dx / math.sqrt(math.pi)
    Vn = math.hypot(U, W)
    Re = rho_g * Vn * L / s.mu_g
    cf = 0.027 / Re ** (1.0 / 7.0)
    # plate: stagnation pressure of the normal velocity U, friction of W
    plate = np.array([0.5 * rho_g * U * U * A / m, 0.5 * rho_g * W * W * cf * A / m])
    # sphere drag of a droplet of diameter L, cross-section pi L^2 / 4
    cd = (1.0 + 0.197 * Re ** 0.63 + 2.6e-4 * Re ** 1.38) * 24.0 / (Re + 1e-5)
    drop = 0.5 * rho_g * Vn * np.array([U, W]) * math.pi * L * L / 4 * cd / m
    expect = frac[:, None] * plate[None, :] + (1 - frac[:, None]) * drop[None, :]
    assert np.allclose(af[rear][:, :2], expect, rtol=1e-9, atol=0)


# ------------------------------------ the reference's analytic validations

def _validation_profiles():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "validation_profiles.json")) as f:
        return json.load(f)


def test_couette_startup_matches_the_reference_profile(po):
    """Whole oracle run of the Couette example to t = 5 s against the
    analytic start-up profile of Examples/Couette/SteadyState.dat (walls
    no-slip; RMS 0.1% of the wall speed measured): viscosity, the moving
    wall's velocity, no-slip walls and the frame clock together."""
    from tests.test_gpu_validation import _profile_error, run_frames
    prof = _validation_profiles()["couette"]["t=5s"]
    case = cases.couette_2d()
    for b in case.blocks[:2]:
        b["no_slip"] = 1
    case.settings.frame_time_interval = 1.0
    o = po.Oracle(case, nthreads=8)
    o.integrate_no_update()
    x, v, _ = run_frames(o, 5.0, 1.0)[5.0]
    rms, mx, _ = _profile_error(x, v, case.n_bound, np.array(prof["z"]), np.array(prof["v"]), 1e-3)
    assert rms <= 5e-3 and mx <= 1e-2


def test_poiseuille_startup_matches_the_reference_profile(po):
    """Whole oracle run of the body-force-driven channel to t = 0.02 s
    against Examples/Poiseuille/SteadyState.dat (RMS 1.6% of the peak
    measured; the profile is young, its values near the wall small)."""
    from tests.test_gpu_validation import _profile_error, run_frames
    prof = _validation_profiles()["poiseuille"]["t=0.02s"]
    case = cases.poiseuille_2d()
    for b in case.blocks[:2]:
        b["no_slip"] = 1
    case.settings.frame_time_interval = 0.02
    o = po.Oracle(case, nthreads=8)
    o.integrate_no_update()
    x, v, _ = run_frames(o, 0.02, 0.02)[0.02]
    u = np.array(prof["v"])
    rms, mx, _ = _profile_error(x, v, case.n_bound, np.array(prof["z"]), u, np.abs(u).max())
    assert rms <= 3e-2 and mx <= 6e-2


# ------------------------------------------------------------- FirstCell

@pytest.mark.parametrize("wall", [False, True], ids=["far_field", "wall"])
def test_first_cell_reflects_or_deletes(po, wall):
    """FirstCell's branch (src/Containment.cpp:466-573) for points just
    under a structured mesh, each crossing the face above it once: across a
    far-field face the point is marked for deletion and left as it was;
    across a wall face it is mirrored, v -> v - 2 (v.n) n with the face
    normal along y, and moved onto the face's plane y = 0.54 dx."""
    from tests.test_gpu_jicf import _first_cell_case
    case = _first_cell_case(wall)
    o = po.Oracle(case, nthreads=2)
    dx = case.settings.particle_step
    rng = np.random.default_rng(4)
    # column positions of the jet, just under the mesh
    cols = case.x[case.b == abi.BACK][:, [0, 2]]
    for k in range(20):
        xz = cols[rng.integers(cols.shape[0])]
        x = np.array([xz[0], (0.54 - rng.uniform(0.005, 0.04)) * dx, xz[1]])
        v = np.array(rng.uniform(-1, 1, 3)) + np.array([0.0, 5.0, 0.0])
        x2, v2, cross, dele = o.first_cell_boundary(x, v)
        assert cross == 1
        if wall:
            assert not dele
            assert abs(x2[1] - 0.54 * dx) <= 1e-15 and x2[0] == x[0] and x2[2] == x[2]
            assert v2[1] == -v[1] and v2[0] == v[0] and v2[2] == v[2]
        else:
            assert dele
            assert np.array_equal(x2, x) and np.array_equal(v2, v)
