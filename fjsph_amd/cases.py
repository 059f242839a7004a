"""Seeded initial conditions for the benchmark and parity workloads.

Each builder returns a Case: host-order arrays (boundary particles first, then
fluid, as Init_Particles orders them, src/Init.cpp:288-484), the settings and
the block table.  Lattices follow the reference's conventions: a cubic
lattice clipped to a sphere for droplets (src/shapes/circle.cpp:163-189),
square/cube blocks indexed i fastest (src/shapes/square.cpp:112-138) and
layered walls (src/shapes/line.cpp:140-154).  Choices the reference leaves to
the user (wall placement, jitter, unstated constants) are stated per builder.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class Case:
    name: str
    settings: abi.Settings
    x: np.ndarray
    v: np.ndarray
    rho: np.ndarray
    p: np.ndarray
    m: np.ndarray
    b: np.ndarray
    n_bound: int
    blocks: list = field(default_factory=list)
    notes: str = ""
    mesh: object = None
    planes: dict = None       # block -> keyword arguments of abi.make_planes
    inlets: list = None       # (block, back host indices, buffer host indices [n_cols][n_buf])
    max_points: int = None

    @property
    def n(self):
        return int(self.x.shape[0])

    @property
    def dim(self):
        return int(self.settings.dim)


def _mass(s):
    return s.rho_rest * s.particle_step ** s.dim


def _finish(name, s, xb, xf, vf, notes, bound_solver=abi.PRESSURE_G, fluid_type=abi.FREE):
    dim = s.dim
    xb = np.asarray(xb, dtype=np.float64).reshape(-1, dim)
    xf = np.asarray(xf, dtype=np.float64).reshape(-1, dim)
    nb, nf = xb.shape[0], xf.shape[0]
    x = np.concatenate([xb, xf], axis=0)
    v = np.zeros_like(x)
    v[nb:] = np.asarray(vf, dtype=np.float64).reshape(-1, dim) if np.ndim(vf) > 1 else np.asarray(vf, dtype=np.float64)
    rho = np.full(nb + nf, s.rho_rest)
    p = np.zeros(nb + nf)
    m = np.full(nb + nf, _mass(s))
    b = np.full(nb + nf, fluid_type, dtype=np.int32)
    b[:nb] = abi.BOUND
    blocks = []
    if nb:
        blocks.append(dict(first=0, last=nb, is_fluid=0, bound_solver=bound_solver))
    blocks.append(dict(first=nb, last=nb + nf, is_fluid=1))
    return Case(name, s, np.ascontiguousarray(x), np.ascontiguousarray(v), rho, p, m, b, nb, blocks, notes)


def dam_break_2d(dx=0.02, precision=abi.PREC_FP64, width=1.0, height=2.0, tank=(4.0, 3.0), layers=3):
    """Plain 2D WCSPH dam break: a 1.0 x 2.0 m column in a 4.0 x 3.0 m tank.

    dx = 0.02 gives 50 x 100 = 5,000 fluid particles.  Wendland C2 with
    H = 1.3 dx, Cole EOS with rho0 = 1000 and c0 = 10 sqrt(2 g H_col),
    delta = 0.1, alpha = 0.1, g = -9.81, mu = sigma = 0 (as in
    Examples/Dam_2D/para), Adami boundary pressure (Examples/Dam_2D/
    boundary.bmap:9), fixed CFL 0.3.  Fluid particles sit at cell centres
    (i + 1/2) dx inside [0, W] x [0, H]; the walls are `layers` rows of
    particles outside the tank at -(k + 1/2) dx, corners included.  No
    jitter (the reference's own jitter is below 1e-17 dx).
    """
    s = abi.default_settings(2)
    s.particle_step = dx
    s.H_fac = 1.3
    s.rho_rest = 1000.0
    s.speed_sound = 10.0 * math.sqrt(2.0 * 9.81 * height)
    s.mu = 0.0
    s.sig = 0.0
    s.visc_alpha = 0.1
    s.dsph_delta = 0.1
    s.cfl = s.cfl_min = s.cfl_max = 0.3
    s.precision = precision
    s.aero_case = abi.NO_AERO
    nx, ny = int(round(width / dx)), int(round(height / dx))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xf = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], axis=1)
    tw, th = tank
    ntx = int(round(tw / dx))
    nty = int(round(th / dx))
    walls = []
    for k in range(layers):
        off = (k + 0.5) * dx
        # bottom row spans the corners
        xs = (np.arange(-layers, ntx + layers) + 0.5) * dx
        walls.append(np.stack([xs, np.full_like(xs, -off)], axis=1))
        ys = (np.arange(0, nty) + 0.5) * dx
        walls.append(np.stack([np.full_like(ys, -off), ys], axis=1))
        walls.append(np.stack([np.full_like(ys, tw + off), ys], axis=1))
    xb = np.concatenate(walls, axis=0)
    return _finish("dam_break_2d", s, xb, xf, 0.0,
                   f"2D dam break dx={dx}: {xf.shape[0]} fluid + {xb.shape[0]} boundary")


def _sphere_lattice(radius, dx, centre, rng, jitter):
    n = int(math.ceil(2.0 * radius / dx))
    n = max(n, 1)
    g = np.arange(n, dtype=np.float64) * dx
    k, j, i = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    pts += np.asarray(centre) - radius
    if jitter:
        pts += rng.uniform(-jitter * dx, jitter * dx, size=pts.shape)
    keep = ((pts - np.asarray(centre)) ** 2).sum(axis=1) <= radius * radius
    return pts[keep]


def _gas(s, speed, rho_g):
    """Constant-velocity gas of speed `speed` along +x with density rho_g:
    p_ref is set so that gamma p / sos^2 = rho_g (the density the Gissler
    force uses, src/Aero.h:97)."""
    s.v_inf[0], s.v_inf[1], s.v_inf[2] = speed, 0.0, 0.0
    s.rho_g = rho_g
    sos2 = s.temp_g * s.R_g * s.gamma_g
    s.p_ref = rho_g * sos2 / s.gamma_g


def droplet_3d(diameter=4e-3, dx=5e-5, seed=1234, precision=abi.PREC_MIXED, speed=20.0, rho_g=1.2, H_fac=2.0):
    """3D single droplet in uniform crossflow.

    A sphere of the given diameter on a cubic lattice of spacing dx (clipped as
    in src/shapes/circle.cpp:163-189), jittered by a seeded U(-0.01dx, +0.01dx)
    per component (the reference jitters by U(0, eps dx), src/shapes/
    square.cpp:103: a deliberate input change).  D = 4 mm, dx = 0.05 mm gives
    about 268k particles.  rho0 = 1000, sigma = 0.0728 with pairwise surface
    tension, Gissler aero in a 20 m/s stream of density 1.2.  Unstated
    constants keep the reference defaults: H_fac = 2, c0 = 300, mu = 8.94e-4,
    alpha = 0.1, delta = 0.1, gravity (0, 0, -9.81).
    """
    s = abi.default_settings(3)
    s.particle_step = dx
    s.H_fac = H_fac
    s.rho_rest = 1000.0
    s.sig = 0.0728
    s.precision = precision
    s.aero_case = abi.GISSLER
    _gas(s, speed, rho_g)
    rng = np.random.default_rng(seed)
    xf = _sphere_lattice(diameter / 2.0, dx, (0.0, 0.0, 0.0), rng, 0.01)
    return _finish("droplet_3d", s, np.zeros((0, 3)), xf, 0.0,
                   f"3D droplet D={diameter} dx={dx}: {xf.shape[0]} particles, seed {seed}")


def droplet_2d(diameter=4e-3, dx=1e-4, centre=(0.0, 0.0), seed=1234, precision=abi.PREC_FP64, speed=20.0,
               rho_g=1.2, H_fac=1.3):
    """2D droplet: a disk of the given diameter on a square lattice of
    spacing dx (jittered by U(-0.01dx, +0.01dx) per component), FREE
    particles, Gissler aero in a constant stream along +x (or a mesh, with
    aero_source set to MESH_INFL).  rho0 = 1000, sigma = 0.0728."""
    s = abi.default_settings(2)
    s.particle_step = dx
    s.H_fac = H_fac
    s.rho_rest = 1000.0
    s.sig = 0.0728
    s.precision = precision
    s.aero_case = abi.GISSLER
    _gas(s, speed, rho_g)
    rng = np.random.default_rng(seed)
    r = diameter / 2.0
    n = max(1, int(math.ceil(2.0 * r / dx)))
    g = np.arange(n, dtype=np.float64) * dx
    j, i = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([i.ravel(), j.ravel()], axis=1) + np.asarray(centre, np.float64) - r
    pts += rng.uniform(-0.01 * dx, 0.01 * dx, size=pts.shape)
    xf = pts[((pts - np.asarray(centre)) ** 2).sum(axis=1) <= r * r]
    return _finish("droplet_2d", s, np.zeros((0, 2)), xf, 0.0,
                   f"2D droplet D={diameter} dx={dx}: {xf.shape[0]} particles, seed {seed}")


def jet_column_3d(n_target=4_000_000, diameter=2e-3, dx=4e-5, seed=7, precision=abi.PREC_MIXED,
                  jet_speed=10.0, coflow=30.0, rho_g=1.2, ale=1, H_fac=2.0):
    """Large 3D coflow jet column: the 4M-particle single-GPU workload.

    A liquid column of the given diameter on a cubic lattice of spacing dx,
    long enough to hold about n_target particles, moving at jet_speed along
    +x inside a constant coflow of `coflow` m/s (relative gas speed
    coflow - jet_speed).  ALE shifting on; Gissler aero from the constant
    coflow velocity.  Jitter U(-0.01dx, 0.01dx), seeded.  Unstated constants
    keep the reference defaults (H_fac = 2, c0 = 300, mu = 8.94e-4,
    alpha = 0.1, delta = 0.1); gravity is off along the jet axis only in the
    sense that the reference default (0, 0, -9.81) is kept.
    """
    s = abi.default_settings(3)
    s.particle_step = dx
    s.H_fac = H_fac
    s.rho_rest = 1000.0
    s.sig = 0.0728
    s.precision = precision
    s.aero_case = abi.GISSLER
    s.ale = ale
    _gas(s, coflow, rho_g)
    r = diameter / 2.0
    nr = int(math.ceil(2.0 * r / dx))
    g = (np.arange(nr) * dx) - r
    yy, zz = np.meshgrid(g, g, indexing="ij")
    keep = (yy ** 2 + zz ** 2) <= r * r
    yz = np.stack([yy[keep], zz[keep]], axis=1)
    per_layer = yz.shape[0]
    nlayers = max(1, int(round(n_target / per_layer)))
    xs = np.arange(nlayers, dtype=np.float64) * dx
    X = np.repeat(xs, per_layer)
    YZ = np.tile(yz, (nlayers, 1))
    pts = np.concatenate([X[:, None], YZ], axis=1)
    rng = np.random.default_rng(seed)
    pts += rng.uniform(-0.01 * dx, 0.01 * dx, size=pts.shape)
    vel = np.zeros_like(pts)
    vel[:, 0] = jet_speed
    return _finish("jet_column_3d", s, np.zeros((0, 3)), pts, vel,
                   f"3D coflow jet column D={diameter} dx={dx}: {pts.shape[0]} particles "
                   f"({nlayers} layers of {per_layer}), seed {seed}")


def pipe_inlet_3d(diameter=6e-4, dx=5e-5, speed=5.0, n_pipe=6, n_buf=4, mode=0, precision=abi.PREC_FP64,
                  H_fac=1.3, gas=20.0, rho_g=1.2, back_to_insert=0.1, exit_to_delete=0.05, max_points=None):
    """Pipe inlet along +z (the inlet zone of the jet-in-crossflow
    configuration, src/shapes/inlet.cpp): a disk lattice of spacing dx
    (columns), n_pipe layers of PIPE particles at z = dx .. n_pipe dx, the
    BACK layer at z = 0 and n_buf BUFFER layers behind it, all moving at
    `speed` along +z.  The insertion plane lies back_to_insert dx above the
    back layer (the first column advance comes after a few steps); the aero
    plane just below the top pipe layer (it turns FREE at the first
    Check_Pipe_Outlet); the deletion plane exit_to_delete dx above it (that
    layer is deleted after a few steps).  mode: 0 fixed-velocity buffer, 1
    dynamic.  Host order: pipe layers, back layer, buffer layers by depth."""
    s = abi.default_settings(3)
    s.particle_step = dx
    s.H_fac = H_fac
    s.rho_rest = 1000.0
    s.sig = 0.0728
    s.precision = precision
    s.aero_case = abi.GISSLER
    _gas(s, gas, rho_g)
    r = diameter / 2.0
    nr = int(math.ceil(2.0 * r / dx)) + 1
    g = (np.arange(nr) - (nr - 1) / 2.0) * dx
    xx, yy = np.meshgrid(g, g, indexing="ij")
    keep = (xx ** 2 + yy ** 2) <= r * r + 1e-12
    disk = np.stack([xx[keep], yy[keep]], axis=1)
    ncol = disk.shape[0]
    def layer(z):
        return np.concatenate([disk, np.full((ncol, 1), z)], axis=1)
    pipe = np.concatenate([layer(k * dx) for k in range(1, n_pipe + 1)])
    back = layer(0.0)
    buf = np.concatenate([layer(-k * dx) for k in range(1, n_buf + 1)])
    pts = np.concatenate([pipe, back, buf])
    vel = np.zeros_like(pts)
    vel[:, 2] = speed
    c = _finish("pipe_inlet_3d", s, np.zeros((0, 3)), pts, vel,
                f"pipe inlet D={diameter} dx={dx}: {ncol} columns, {n_pipe} pipe layers, {n_buf} buffer layers, "
                f"{'dynamic' if mode else 'fixed-velocity'} buffer", fluid_type=abi.PIPE)
    npipe = pipe.shape[0]
    c.b[npipe:npipe + ncol] = abi.BACK
    c.b[npipe + ncol:] = abi.BUFFER
    back_ids = np.arange(npipe, npipe + ncol)
    buffer_ids = np.stack([npipe + ncol + k * ncol + np.arange(ncol) for k in range(n_buf)], axis=1)
    top = n_pipe * dx
    c.planes = {0: dict(inlet=1, fixed_vel_or_dynamic=mode,
                        insert=((0.0, 0.0, 1.0), back_to_insert * dx),
                        aero=((0.0, 0.0, 1.0), top - 0.1 * dx),
                        delete=((0.0, 0.0, 1.0), top + exit_to_delete * dx))}
    c.inlets = [(0, back_ids, buffer_ids)]
    c.max_points = max_points
    return c


def _wall_rows(x0, x1, y, dx, layers, outward):
    """`layers` rows of wall particles at y + outward (k + 1/2) dx, k <
    layers, spaced dx from x0 + dx/2 while below x1."""
    xs = x0 + (np.arange(int(round((x1 - x0) / dx))) + 0.5) * dx
    return np.concatenate([np.stack([xs, np.full_like(xs, y + outward * (k + 0.5) * dx)], axis=1)
                           for k in range(layers)])


def _validation_case(name, s, walls, xf, notes):
    """Walls (list of (points, spec)) first, one block each, then the fluid."""
    xb = np.concatenate([w for w, _ in walls]) if walls else np.zeros((0, 2))
    c = _finish(name, s, xb, xf, 0.0, notes)
    blocks, first = [], 0
    for pts, spec in walls:
        blocks.append(dict(first=first, last=first + pts.shape[0], is_fluid=0, **spec))
        first += pts.shape[0]
    blocks.append(dict(first=first, last=c.n, is_fluid=1))
    c.blocks = blocks
    return c


def couette_2d(dx=0.02, length=3.0, height=1.0, v_wall=1e-3, layers=4, precision=abi.PREC_FP64):
    """Couette start-up (Examples/Couette: para, fluid.bmap, boundary.bmap):
    rho0 = 1, mu = 1e-2, sigma = 0, c0 = 1, alpha = 0.01, H = 2 dx, CFL 1,
    maximum step 0.5 s, no gravity.  Fluid in 0 < x < length between a wall
    at rest below y = 0 and a wall above y = height whose particles move at
    v_wall along +x (the block velocity, which enters the fluid's viscous
    terms; wall positions stay, as in the reference); 4 rows of wall
    particles each (Adami pressure, free slip), reaching 4 dx past the
    fluid's ends, which are free.  The example's lattice puts its first wall
    row on y = 0 and its fluid rows from dx; here the fluid sits at cell
    centres ((i + 1/2) dx, (j + 1/2) dx) and the wall rows at -(k + 1/2) dx
    and height + (k + 1/2) dx, so both walls lie where the analytic solution
    puts them.  The example's analytic profiles are in
    Examples/Couette/SteadyState.dat."""
    s = abi.default_settings(2)
    s.particle_step = dx
    s.H_fac = 2.0
    s.rho_rest = 1.0
    s.mu = 1e-2
    s.sig = 0.0
    s.speed_sound = 1.0
    s.visc_alpha = 0.01
    s.cfl = 1.0
    s.delta_t_max = 0.5
    s.grav[0] = s.grav[1] = 0.0
    s.precision = precision
    s.aero_case = abi.NO_AERO
    nx, ny = int(round(length / dx)), int(round(height / dx))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xf = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], axis=1)
    pad = 4 * dx
    bottom = _wall_rows(-pad, length + pad, 0.0, dx, layers, -1.0)
    top = _wall_rows(-pad, length + pad, height, dx, layers, +1.0)
    return _validation_case("couette_2d", s, [(bottom, dict(bound_solver=abi.PRESSURE_G)),
                                              (top, dict(bound_solver=abi.PRESSURE_G, vel=(v_wall, 0.0, 0.0)))],
                            xf, f"Couette {nx}x{ny} fluid, dx={dx}, top wall at {v_wall} m/s")


def poiseuille_2d(dx=4e-5, length=5e-3, half_height=5e-4, g=1e-4, layers=4, precision=abi.PREC_FP64):
    """Body-force-driven channel start-up (Examples/Poiseuille: para,
    boundary.bmap): rho0 = 1000, mu = 1e-3, sigma = 0, c0 = 1, alpha =
    0.001, H = 2 dx, CFL 1.111112, gravity g along +x, walls at y = +-0.5 mm
    of 4 particle rows (Adami pressure, free slip).  The example feeds the
    channel from a dynamic inlet and drains it past x = 5 mm; here the 5 mm
    channel's ends are free, so its core accelerates under the body force
    alone, as the fully developed flow the analytic profiles describe does
    (the slug moves about 6 um in 0.5 s, against a spacing of 40 um).
    Lattice: fluid at cell centres, wall rows at -+(h + (k + 1/2) dx).  The
    analytic profiles are in Examples/Poiseuille/SteadyState.dat."""
    s = abi.default_settings(2)
    s.particle_step = dx
    s.H_fac = 2.0
    s.rho_rest = 1000.0
    s.mu = 1e-3
    s.sig = 0.0
    s.speed_sound = 1.0
    s.visc_alpha = 0.001
    s.cfl = 1.111112
    s.grav[0], s.grav[1] = g, 0.0
    s.precision = precision
    s.aero_case = abi.NO_AERO
    nx, ny = int(round(length / dx)), int(round(2 * half_height / dx))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xf = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx - half_height], axis=1)
    pad = 4 * dx
    bottom = _wall_rows(-pad, length + pad, -half_height, dx, layers, -1.0)
    top = _wall_rows(-pad, length + pad, half_height, dx, layers, +1.0)
    return _validation_case("poiseuille_2d", s, [(bottom, dict(bound_solver=abi.PRESSURE_G)),
                                                 (top, dict(bound_solver=abi.PRESSURE_G))],
                            xf, f"Poiseuille channel {nx}x{ny} fluid, dx={dx}")


def standing_column_2d(dx=0.01, width=1.0, height=2.0, layers=4, precision=abi.PREC_FP64):
    """Hydrostatic column (Examples/Standing_Column: para, fluid.bmap,
    boundary.bmap): a 1 x 2 m water column in a tank of its width, rho0 =
    1000, mu = 8.94e-4, sigma = 0, c0 = 100, alpha = 0.1, H = 2 dx, CFL 1,
    g = -9.81 along y, walls of 4 particle rows (Adami pressure) left, right
    and below, and the example's hydrostatic initial pressure p = rho0 g (H
    - y) (density from the Cole equation of state).  Fluid at cell centres,
    wall rows outside the tank at -(k + 1/2) dx.  The ideal line P / (rho g
    H) = (H - Y) / H is Examples/Standing_Column/Ideal.dat."""
    s = abi.default_settings(2)
    s.particle_step = dx
    s.H_fac = 2.0
    s.rho_rest = 1000.0
    s.mu = 8.94e-4
    s.sig = 0.0
    s.speed_sound = 100.0
    s.visc_alpha = 0.1
    s.cfl = 1.0
    s.grav[0], s.grav[1] = 0.0, -9.81
    s.precision = precision
    s.aero_case = abi.NO_AERO
    nx, ny = int(round(width / dx)), int(round(height / dx))
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    xf = np.stack([(i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx], axis=1)
    walls = []
    for k in range(layers):
        off = (k + 0.5) * dx
        xs = (np.arange(-layers, nx + layers) + 0.5) * dx
        walls.append(np.stack([xs, np.full_like(xs, -off)], axis=1))
        ys = (np.arange(0, ny + 4) + 0.5) * dx
        walls.append(np.stack([np.full_like(ys, -off), ys], axis=1))
        walls.append(np.stack([np.full_like(ys, width + off), ys], axis=1))
    xb = np.concatenate(walls)
    c = _validation_case("standing_column_2d", s, [(xb, dict(bound_solver=abi.PRESSURE_G))], xf,
                         f"standing column {nx}x{ny}, dx={dx}")
    nb = c.n_bound
    B = s.rho_rest * s.speed_sound ** 2 / s.gam
    p = s.rho_rest * 9.81 * (height - c.x[nb:, 1])
    c.p = c.p.copy()
    c.rho = c.rho.copy()
    c.p[nb:] = p
    c.rho[nb:] = s.rho_rest * (1.0 + p / B) ** (1.0 / s.gam)
    c.m = np.full(c.n, s.rho_rest * dx * dx)
    return c


def hollow_cylinder_3d(radius, length, dx, layers):
    """Wall of a pipe along +y ending at y = 0 (src/shapes/cylinder.cpp
    create_hollow_cylinder, square ordering, rotated so that its axis is
    +y): rings of ni = ceil(2 pi radius / dx) particles, dtheta = 2 pi / ni,
    at radii radius + k dx for k < layers, every dx from y = 0 down to
    y = -(nj - 1) dx, nj = ceil(length / dx) + 1."""
    ni = max(1, int(math.ceil(2.0 * math.pi / (dx / radius))))
    nj = max(1, int(math.ceil(length / dx)) + 1)
    dtheta = 2.0 * math.pi / ni
    pts = []
    for jj in range(nj):
        for kk in range(layers):
            r = radius + kk * dx
            th = np.arange(ni) * dtheta
            pts.append(np.stack([np.cos(th) * r, np.full(ni, -jj * dx), np.sin(th) * r], axis=1))
    return np.concatenate(pts)


def jet_in_crossflow_3d(n_target=1_000_000, diameter=1e-3, dx=3e-5, speed=5.0, n_pipe=8, n_buf=4, mode=0,
                        wall_layers=2, precision=abi.PREC_MIXED, H_fac=2.0, gas=50.0, rho_g=1.2262,
                        jet_layers=None, back_to_insert=0.1, delete_above=None):
    """3D jet in crossflow: a liquid jet of diameter D issuing at `speed`
    along +y from a pipe into a crossflow along +x (Examples/Crossflow:
    para3D, fluid_3D.bmap, pipe_3D.bmap).

    - pipe wall: a hollow cylinder of ghost particles (pipe_3D.bmap: Shape
      Cylinder, Sub-shape Hollow, Boundary solver Ghost, no slip 0, two
      radial particle layers) of inner radius D/2 + dx, from the pipe exit
      at y = 0 down past the inlet buffer;
    - inlet (fluid_3D.bmap: Shape Inlet, Sub-shape Circle): columns on the
      square lattice of spacing dx inside the circle of radius D/2, n_pipe
      PIPE layers up to the exit (y = 0), the BACK layer below them and
      n_buf BUFFER layers below that, all at `speed` along +y; the
      insertion plane back_to_insert dx above the back layer, the
      aerodynamic entry plane half a spacing above the exit (PIPE particles
      beyond it turn FREE), mode 0 fixed-velocity / 1 dynamic buffer;
    - the jet already issued: jet_layers FREE layers of the columns above
      the exit at `speed` (by default as many as make n_target particles in
      all: the configuration grows to 1M particles, and the state starts
      grown);
    - a deletion plane `delete_above` (m, along +y) if given.
    Physics of para3D: rho0 = 997, c0 = 300, alpha = 0.05, mu = 8.94e-4,
    sigma = 0.0708, Gissler drag, gas rho_g = 1.2262 in a stream of `gas`
    m/s along +x (constant; cases.with_mesh gives it a mesh).  The
    reference example's fluid radius is 4e-4 in a wall of radius 4.9e-4 at
    dx = 3e-5; here the wall sits one spacing outside the outermost column.
    Host order: wall, jet layers (top first), pipe layers (top first), back
    layer, buffer layers by depth: the particles a deletion plane removes
    then come before the inlet's, as in the reference's own runs (it inserts
    at the block's end and shifts the inlet's index lists by every deletion
    in the block, src/Integration.cpp:186-205)."""
    s = abi.default_settings(3)
    s.particle_step = dx
    s.H_fac = H_fac
    s.rho_rest = 997.0
    s.speed_sound = 300.0
    s.visc_alpha = 0.05
    s.mu = 8.94e-4
    s.sig = 0.0708
    s.precision = precision
    s.aero_case = abi.GISSLER
    s.mu_g = 1.85e-5
    _gas(s, gas, rho_g)
    r = diameter / 2.0
    nr = int(math.ceil(2.0 * r / dx)) + 1
    g = (np.arange(nr) - (nr - 1) / 2.0) * dx
    xx, zz = np.meshgrid(g, g, indexing="ij")
    keep = (xx ** 2 + zz ** 2) <= r * r + 1e-12
    disk = np.stack([xx[keep], zz[keep]], axis=1)
    ncol = disk.shape[0]

    def layer(y):
        return np.stack([disk[:, 0], np.full(ncol, y), disk[:, 1]], axis=1)

    wall = hollow_cylinder_3d(r + dx, (n_pipe + n_buf + 2) * dx, dx, wall_layers)
    if jet_layers is None:
        jet_layers = max(0, int(round((n_target - wall.shape[0]) / ncol)) - (n_pipe + 1 + n_buf))
    jet = (np.concatenate([layer(k * dx) for k in range(jet_layers, 0, -1)]) if jet_layers
           else np.zeros((0, 3)))
    pipe = np.concatenate([layer(-k * dx) for k in range(n_pipe)])
    y_back = -n_pipe * dx
    back = layer(y_back)
    buf = np.concatenate([layer(y_back - k * dx) for k in range(1, n_buf + 1)])
    pts = np.concatenate([jet, pipe, back, buf])
    vel = np.zeros_like(pts)
    vel[:, 1] = speed
    c = _finish("jet_in_crossflow_3d", s, wall, pts, vel,
                f"jet in crossflow D={diameter} dx={dx}: {ncol} columns, {wall.shape[0]} ghost wall particles, "
                f"{n_pipe} pipe + {n_buf} buffer + {jet_layers} jet layers", bound_solver=abi.GHOST,
                fluid_type=abi.PIPE)
    nb = wall.shape[0]
    nj, npipe = jet.shape[0], pipe.shape[0]
    c.b[nb:nb + nj] = abi.FREE
    c.b[nb + nj + npipe:nb + nj + npipe + ncol] = abi.BACK
    c.b[nb + nj + npipe + ncol:] = abi.BUFFER
    back_ids = np.arange(nb + nj + npipe, nb + nj + npipe + ncol)
    buffer_ids = np.stack([nb + nj + npipe + ncol + k * ncol + np.arange(ncol) for k in range(n_buf)], axis=1)
    planes = dict(inlet=1, fixed_vel_or_dynamic=mode,
                  insert=((0.0, 1.0, 0.0), y_back + back_to_insert * dx),
                  aero=((0.0, 1.0, 0.0), 0.5 * dx))
    if delete_above is not None:
        planes["delete"] = ((0.0, 1.0, 0.0), float(delete_above))
    c.planes = {1: planes}
    c.inlets = [(1, back_ids, buffer_ids)]
    c.jet_top = jet_layers * dx
    c.n_columns = ncol
    return c


def fluid_block(dim, n_side, dx=1e-3, jitter=0.0, seed=0, precision=abi.PREC_FP64, H_fac=1.3,
                with_walls=False, layers=3, speed=0.0, aero=abi.NO_AERO, sig=0.0, mu=8.94e-4):
    """Small square/cube of fluid (i fastest, src/shapes/square.cpp:112-138),
    optionally in an open box of `layers` wall rows: the unit-test workload."""
    s = abi.default_settings(dim)
    s.particle_step = dx
    s.H_fac = H_fac
    s.precision = precision
    s.sig = sig
    s.mu = mu
    s.aero_case = aero
    if aero != abi.NO_AERO:
        _gas(s, speed, 1.2)
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * dx
    mesh = np.meshgrid(*([g] * dim), indexing="ij")
    xf = np.stack([m.ravel() for m in reversed(mesh)], axis=1)
    if jitter:
        xf += rng.uniform(-jitter * dx, jitter * dx, size=xf.shape)
    xb = np.zeros((0, dim))
    if with_walls:
        L = n_side * dx
        walls = []
        span = (np.arange(-layers, n_side + layers) + 0.5) * dx
        for k in range(layers):
            off = (k + 0.5) * dx
            if dim == 2:
                walls.append(np.stack([span, np.full_like(span, -off)], axis=1))
                ys = (np.arange(n_side + layers) + 0.5) * dx
                walls.append(np.stack([np.full_like(ys, -off), ys], axis=1))
                walls.append(np.stack([np.full_like(ys, L + off), ys], axis=1))
            else:
                a, bb = np.meshgrid(span, span, indexing="ij")
                walls.append(np.stack([a.ravel(), bb.ravel(), np.full(a.size, -off)], axis=1))
        xb = np.concatenate(walls, axis=0)
    return _finish(f"block_{dim}d", s, xb, xf, 0.0, f"{dim}D block {n_side}^{dim}")


def irregular_cloud(dim, H_fac=1.3, dx=1e-3, seed=0, precision=abi.PREC_FP64):
    """A block jittered by up to half a spacing per component beside a denser
    Gaussian blob: cells of uneven counts and no lattice order (the test
    workload of the neighbour search on irregular geometry)."""
    n_side = 40 if dim == 2 else 18
    blk = fluid_block(dim, n_side, dx=dx, jitter=0.5, seed=seed, H_fac=H_fac, precision=precision)
    rng = np.random.default_rng(seed + 1)
    n_blob = 1500 if dim == 2 else 2500
    blob = rng.normal(0.0, 3.0 * dx, size=(n_blob, dim)) + blk.x.max(axis=0)
    xf = np.concatenate([blk.x, blob], axis=0)
    return _finish(f"cloud_{dim}d", blk.settings, np.zeros((0, dim)), xf, 0.0,
                   f"{dim}D jittered block + Gaussian blob: {xf.shape[0]} particles, seed {seed}")


def batched_droplets(n_drops=256, dx=5e-5, seed=99, precision=abi.PREC_MIXED, d_range=(0.5e-3, 2e-3),
                     u_range=(5.0, 40.0), spacing_factor=1.5):
    """Many independent 3D droplets of uneven size in one particle set.

    The stated radius range [0.5, 2] mm at dx = 0.05 mm gives 4k-268k
    particles per drop, not the stated 2k-60k; this builder treats the range
    as DIAMETERS (log-uniform), which gives about 0.5k-33k particles per
    drop.  Drops are laid out on a grid with gaps of spacing_factor x the
    largest diameter so no two interact.  Each drop gets its own crossflow
    speed U[5, 40] m/s through its initial velocity relative to a common gas
    at rest (the relative velocity is what the Gissler force sees).
    """
    s = abi.default_settings(3)
    s.particle_step = dx
    s.rho_rest = 1000.0
    s.sig = 0.0728
    s.precision = precision
    s.aero_case = abi.GISSLER
    _gas(s, 0.0, 1.2)
    rng = np.random.default_rng(seed)
    logd = rng.uniform(math.log(d_range[0]), math.log(d_range[1]), size=n_drops)
    diams = np.exp(logd)
    speeds = rng.uniform(u_range[0], u_range[1], size=n_drops)
    side = int(math.ceil(n_drops ** (1.0 / 3.0)))
    pitch = spacing_factor * d_range[1] + 8 * s.H_fac * dx
    pts, vels, owner = [], [], []
    for k in range(n_drops):
        c = np.array([k % side, (k // side) % side, k // (side * side)], dtype=np.float64) * pitch
        p = _sphere_lattice(diams[k] / 2.0, dx, c, rng, 0.01)
        pts.append(p)
        v = np.zeros_like(p)
        v[:, 0] = -speeds[k]
        vels.append(v)
        owner.append(np.full(p.shape[0], k, dtype=np.int32))
    xf = np.concatenate(pts)
    vf = np.concatenate(vels)
    case = _finish("batched_droplets", s, np.zeros((0, 3)), xf, vf,
                   f"{n_drops} droplets, diameters log-U{d_range}, {xf.shape[0]} particles")
    case.owner = np.concatenate(owner)
    return case


def drop_of_batch(case, g):
    """Drop g of a batched_droplets case as a Case of its own (its particles
    in batch host order, the batch's settings): the standalone simulation
    the batch's group g must follow."""
    sel = np.flatnonzero(case.owner == g)
    sub = _finish(f"drop_{g}", case.settings.copy(), np.zeros((0, case.dim)), case.x[sel], case.v[sel],
                  f"drop {g} of {case.name}: {sel.size} particles")
    sub.rho, sub.p, sub.m, sub.b = case.rho[sel].copy(), case.p[sel].copy(), case.m[sel].copy(), case.b[sel].copy()
    return sub


@dataclass
class Mesh:
    """Aerodynamic mesh in the flat layout of fjsph_set_mesh (MESH of
    src/Var.h:396-451): CSR faces with left/right cells (-2 marks the outer
    boundary), CSR cells, and the gas solution at the vertices."""
    verts: np.ndarray
    face_off: np.ndarray
    face_verts: np.ndarray
    face_lr: np.ndarray
    cell_off: np.ndarray
    cell_faces: np.ndarray
    vert_vel: np.ndarray
    vert_p: np.ndarray
    vert_rho: np.ndarray
    simplices: np.ndarray = None
    tri: object = None  # the scipy triangulation the cells come from (point location for sampling)

    @property
    def nverts(self):
        return self.verts.shape[0]

    @property
    def nfaces(self):
        return self.face_off.size - 1

    @property
    def ncells(self):
        return self.cell_off.size - 1


def cylinder_flow(x, speed, radius, centre, rho_g, p_inf):
    """Potential flow past a circular cylinder (axis z) of the given radius
    and centre in a uniform stream `speed` along +x: velocity, Bernoulli
    pressure, constant density.  Points inside the cylinder get the surface
    value of their direction."""
    dim = x.shape[1]
    dx = x[:, 0] - centre[0]
    dy = x[:, 1] - centre[1]
    r2 = np.maximum(dx * dx + dy * dy, radius * radius)
    a2 = radius * radius
    u = speed * (1.0 - a2 * (dx * dx - dy * dy) / (r2 * r2))
    v = -speed * 2.0 * a2 * dx * dy / (r2 * r2)
    vel = np.zeros((x.shape[0], dim))
    vel[:, 0] = u
    vel[:, 1] = v
    p = p_inf + 0.5 * rho_g * (speed * speed - (u * u + v * v))
    return vel, p, np.full(x.shape[0], rho_g)


def delaunay_mesh(lo, hi, n_points, seed=7, speed=20.0, rho_g=1.2, p_inf=101325.0, cyl_radius=None,
                  cyl_centre=None):
    """Seeded Delaunay mesh (scipy) of random points plus the box corners in
    [lo, hi], tetrahedra (3D) or triangles (2D), with the potential flow of
    `cylinder_flow` at the vertices.  Faces are the simplices' facets, each
    stored once; the first cell seen is left, a facet on the hull gets -2."""
    from itertools import combinations

    from scipy.spatial import Delaunay

    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    dim = lo.size
    rng = np.random.default_rng(seed)
    corners = np.array(np.meshgrid(*[[lo[k], hi[k]] for k in range(dim)], indexing="ij")).reshape(dim, -1).T
    pts = np.concatenate([corners, rng.uniform(lo, hi, size=(n_points, dim))], axis=0)
    tri = Delaunay(pts)
    simp = tri.simplices.astype(np.int64)
    nc = simp.shape[0]
    combos = list(combinations(range(dim + 1), dim))
    facets = np.stack([simp[:, list(cmb)] for cmb in combos], axis=1).reshape(-1, dim)
    key = np.sort(facets, axis=1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    nf = uniq.shape[0]
    owner = np.repeat(np.arange(nc), dim + 1)
    lr = np.full((nf, 2), -2, dtype=np.int64)
    first = np.full(nf, -1, dtype=np.int64)
    order = np.argsort(inv, kind="stable")
    inv_sorted = inv[order]
    starts = np.searchsorted(inv_sorted, np.arange(nf))
    lr[:, 0] = owner[order[starts]]
    second = starts + 1
    has2 = (second < inv_sorted.size) & (inv_sorted[np.minimum(second, inv_sorted.size - 1)] == np.arange(nf))
    lr[has2, 1] = owner[order[second[has2]]]
    del first
    face_off = np.arange(nf + 1, dtype=np.int64) * dim
    face_verts = uniq.reshape(-1).astype(np.int64)
    cell_off = np.arange(nc + 1, dtype=np.int64) * (dim + 1)
    cell_faces = inv.astype(np.int64)
    if cyl_radius is None:
        cyl_radius = 0.25 * float(hi[1] - lo[1])
    if cyl_centre is None:
        cyl_centre = (float(lo[0]) - 1.5 * cyl_radius, 0.5 * float(lo[1] + hi[1]))
    vel, p, rho = cylinder_flow(pts, speed, cyl_radius, cyl_centre, rho_g, p_inf)
    return Mesh(np.ascontiguousarray(pts), face_off, face_verts, np.ascontiguousarray(lr.reshape(-1)), cell_off,
                cell_faces, np.ascontiguousarray(vel), p, rho, simplices=simp, tri=tri)


_KUHN_PERMS = np.array([(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)], dtype=np.int64)


def kuhn_mesh(lo, hi, n_cells, speed=20.0, rho_g=1.2, p_inf=101325.0, cyl_radius=None, cyl_centre=None):
    """Structured 3D tetrahedral mesh of about n_cells cells over [lo, hi]:
    a grid of cubes, each split into the 6 tetrahedra of its main diagonal
    (p0, p0 + e_s0, p0 + e_s0 + e_s1, p0 + (1,1,1) for the permutations s
    of the axes; the split is conforming across cubes).  Faces are
    enumerated directly: per cube the 6 faces through its diagonal, and
    per grid vertex p0 the 2 triangles of the cube face x_d = p0_d that
    starts at p0, for each axis d (the lower face of the cube at p0 and the
    upper face of the cube below it).  The gas solution at the vertices is
    the potential flow of `cylinder_flow`, as for delaunay_mesh."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    ext = hi - lo
    h = (float(np.prod(ext)) * 6.0 / float(n_cells)) ** (1.0 / 3.0)
    n = np.maximum(1, np.round(ext / h)).astype(np.int64)
    nx, ny, nz = (int(v) for v in n)
    vx, vy = nx + 1, ny + 1
    g = [np.linspace(lo[k], hi[k], int(n[k]) + 1) for k in range(3)]
    pts = np.stack(np.meshgrid(g[0], g[1], g[2], indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    pts = np.ascontiguousarray(pts)
    nv = pts.shape[0]
    ncube = nx * ny * nz
    ci = np.arange(ncube, dtype=np.int64)
    cx, cy, cz = ci % nx, (ci // nx) % ny, ci // (nx * ny)
    v0 = (cz * vy + cy) * vx + cx                       # cube origin vertex
    step = np.array([1, vx, vx * vy], dtype=np.int64)   # vertex index step along each axis
    # ---- faces: interior faces of cube c (id 6c + m), lower faces of vertex v (id 6 ncube + 6 v + 2 d + t)
    nint = 6 * ncube
    nfaces_all = nint + 6 * nv
    # ---- cells: id 6c + s
    perms = _KUHN_PERMS
    s0, s1, s2 = perms[:, 0], perms[:, 1], perms[:, 2]
    cell_faces = np.empty((ncube, 6, 4), dtype=np.int64)
    t_omit0 = (s1 > s2).astype(np.int64)   # triangle of the upper face in direction s0
    t_omit3 = (s0 > s1).astype(np.int64)   # triangle of the lower face in direction s2
    cell_faces[:, :, 0] = nint + 6 * (v0[:, None] + step[s0][None, :]) + 2 * s0[None, :] + t_omit0[None, :]
    cell_faces[:, :, 1] = 6 * ci[:, None] + 3 + s2[None, :]
    cell_faces[:, :, 2] = 6 * ci[:, None] + s0[None, :]
    cell_faces[:, :, 3] = nint + 6 * v0[:, None] + 2 * s2[None, :] + t_omit3[None, :]
    cell_faces = cell_faces.reshape(-1, 4)
    simplices = np.empty((ncube, 6, 4), dtype=np.int64)
    simplices[:, :, 0] = v0[:, None]
    simplices[:, :, 1] = v0[:, None] + step[s0][None, :]
    simplices[:, :, 2] = simplices[:, :, 1] + step[s1][None, :]
    simplices[:, :, 3] = v0[:, None] + step.sum()
    simplices = simplices.reshape(-1, 4)
    # ---- face vertices and left/right cells
    fv = np.empty((nfaces_all, 3), dtype=np.int64)
    lr = np.full((nfaces_all, 2), -2, dtype=np.int64)
    m = np.arange(6)
    mid = np.where(m < 3, step[m % 3], step.sum() - step[m % 3])
    fv[:nint, 0] = np.repeat(v0, 6)
    fv[:nint, 1] = (v0[:, None] + mid[None, :]).reshape(-1)
    fv[:nint, 2] = np.repeat(v0 + step.sum(), 6)
    # interior face m < 3: omit-p2 face of the 2 tets with s0 = m; m >= 3: omit-p1 face of the tets with s2 = m - 3
    pair = np.array([np.nonzero(s0 == k)[0] for k in range(3)] + [np.nonzero(s2 == k)[0] for k in range(3)])
    lr[:nint] = (6 * ci[:, None, None] + pair[None, :, :]).reshape(-1, 2)
    vi = np.arange(nv, dtype=np.int64)
    vxi, vyi, vzi = vi % vx, (vi // vx) % vy, vi // (vx * vy)
    vc = np.stack([vxi, vyi, vzi], 1)
    lim = np.array([nx, ny, nz])
    for d in range(3):
        a, b = [k for k in range(3) if k != d]
        for t in range(2):
            ids = nint + 6 * vi + 2 * d + t
            first = a if t == 0 else b
            fv[ids, 0] = vi
            fv[ids, 1] = vi + step[first]
            fv[ids, 2] = vi + step[a] + step[b]
            # lower face of the cube at v (tet with s2 = d, first step `first`)
            s_lo = np.nonzero((s2 == d) & (s0 == first))[0][0]
            in_lo = np.all(vc < lim, axis=1)
            # upper face of the cube at v - e_d (tet with s0 = d, s1 = first)
            s_up = np.nonzero((s0 == d) & (s1 == first))[0][0]
            below = vc.copy()
            below[:, d] -= 1
            in_up = np.all((below >= 0) & (below < lim), axis=1)
            cube_lo = (vzi * ny + vyi) * nx + vxi
            cube_up = (below[:, 2] * ny + below[:, 1]) * nx + below[:, 0]
            c_lo = np.where(in_lo, 6 * cube_lo + s_lo, -2)
            c_up = np.where(in_up, 6 * cube_up + s_up, -2)
            both = in_lo & in_up
            lr[ids, 0] = np.where(both, np.minimum(c_lo, c_up), np.where(in_lo, c_lo, c_up))
            lr[ids, 1] = np.where(both, np.maximum(c_lo, c_up), -2)
    # ---- keep the faces some cell has
    used = np.zeros(nfaces_all, dtype=bool)
    used[cell_faces.reshape(-1)] = True
    remap = np.cumsum(used) - 1
    fv = fv[used]
    lr = lr[used]
    cell_faces = remap[cell_faces]
    nf = fv.shape[0]
    face_off = np.arange(nf + 1, dtype=np.int64) * 3
    cell_off = np.arange(cell_faces.shape[0] + 1, dtype=np.int64) * 4
    if cyl_radius is None:
        cyl_radius = 0.25 * float(hi[1] - lo[1])
    if cyl_centre is None:
        cyl_centre = (float(lo[0]) - 1.5 * cyl_radius, 0.5 * float(lo[1] + hi[1]))
    vel, p, rho = cylinder_flow(pts, speed, cyl_radius, cyl_centre, rho_g, p_inf)
    return Mesh(pts, face_off, np.ascontiguousarray(fv.reshape(-1)), np.ascontiguousarray(lr.reshape(-1)), cell_off,
                np.ascontiguousarray(cell_faces.reshape(-1)), np.ascontiguousarray(vel), p, rho, simplices=simplices)


def quad_mesh(nx, ny, h, vel, p=1.0e5, rho=1.2, lo=(0.0, 0.0)):
    """Structured 2D mesh of nx x ny square cells of side h from `lo`, with a
    uniform gas (velocity `vel`, pressure p, density rho) at the vertices:
    vertical edges first (row by row), then horizontal ones; boundary edges
    have the far-field marker -2 as right cell."""
    xs = lo[0] + np.arange(nx + 1) * h
    ys = lo[1] + np.arange(ny + 1) * h
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    verts = np.stack([X.ravel(), Y.ravel()], axis=1)

    def vid(i, j):
        return j * (nx + 1) + i

    def cid(i, j):
        return j * nx + i

    faces, lr = [], []
    for j in range(ny):
        for i in range(nx + 1):
            faces.append((vid(i, j), vid(i, j + 1)))
            cells = [c for c in ((cid(i - 1, j) if i > 0 else -2), (cid(i, j) if i < nx else -2))]
            lr.append(tuple(sorted(cells, key=lambda c: c < 0)))
    for j in range(ny + 1):
        for i in range(nx):
            faces.append((vid(i, j), vid(i + 1, j)))
            cells = [c for c in ((cid(i, j - 1) if j > 0 else -2), (cid(i, j) if j < ny else -2))]
            lr.append(tuple(sorted(cells, key=lambda c: c < 0)))
    faces = np.array(faces, np.int64)
    lr = np.array(lr, np.int64)
    cf = [[] for _ in range(nx * ny)]
    for f, (l, r) in enumerate(lr):
        cf[l].append(f)
        if r >= 0:
            cf[r].append(f)
    co = np.zeros(nx * ny + 1, np.int64)
    co[1:] = np.cumsum([len(c) for c in cf])
    nv = verts.shape[0]
    return Mesh(verts=verts, face_off=np.arange(0, 2 * len(faces) + 1, 2, dtype=np.int64), face_verts=faces.ravel(),
                face_lr=lr, cell_off=co, cell_faces=np.concatenate(cf).astype(np.int64),
                vert_vel=np.tile(np.asarray(vel, np.float64), (nv, 1)), vert_p=np.full(nv, float(p)),
                vert_rho=np.full(nv, float(rho)))


def with_mesh(case, n_points=20000, margin=4.0, seed=7, speed=None, kind="delaunay", n_cells=None, box=None):
    """Switch a case to the mesh aero source with a mesh around its
    particles (`margin` particle spacings beyond their bounding box, or the
    given box (lo, hi)): a seeded Delaunay mesh of n_points random points
    (kind "delaunay") or a structured tetrahedral mesh of about n_cells
    cells (kind "kuhn", 3D)."""
    s = case.settings
    dx = s.particle_step
    if box is not None:
        lo = np.asarray(box[0], dtype=np.float64)
        hi = np.asarray(box[1], dtype=np.float64)
        if np.any(case.x.min(axis=0) <= lo) or np.any(case.x.max(axis=0) >= hi):
            raise ValueError("with_mesh: the particles must lie inside the box")
    else:
        lo = case.x.min(axis=0) - margin * dx
        hi = case.x.max(axis=0) + margin * dx
    sp = float(speed if speed is not None else s.v_inf[0])
    rho_g = float(s.rho_g)
    p_inf = float(s.p_ref)
    if kind == "kuhn":
        case.mesh = kuhn_mesh(lo, hi, n_cells, speed=sp, rho_g=rho_g, p_inf=p_inf)
    else:
        case.mesh = delaunay_mesh(lo, hi, n_points, seed=seed, speed=sp, rho_g=rho_g, p_inf=p_inf)
    s.aero_source = abi.MESH_INFL
    case.name = case.name + "_mesh"
    return case
