"""Training of the MLP surrogate of the aerodynamic mesh field.

The engine trains the network with its own HIP kernels (BF16 MFMA forward
and backward passes with FP32 accumulation, Adam on FP32 master weights)
and evaluates it in its inference kernel once fjsph_set_surrogate has the
weights; this module draws the samples and sets up the call.  Samples are
points drawn uniformly in the mesh's bounding box, labelled with the
values of the cell that contains them (the same piecewise-constant field
the mesh lookup returns).  The network maps the normalised position to the
normalised (velocity, pressure, density): n_hidden layers of width 64 with
GELU (erf form).
"""
from dataclasses import dataclass

import numpy as np

WIDTH = 64


@dataclass
class Surrogate:
    n_hidden: int
    weights: np.ndarray   # flat FP32: W_1 [64][dim], W_2..W_n [64][64], W_out [dim+2][64]
    biases: np.ndarray    # flat FP32: b_1..b_n [64], b_out [dim+2]
    in_shift: np.ndarray
    in_scale: np.ndarray
    out_shift: np.ndarray
    out_scale: np.ndarray
    train_rmse: np.ndarray = None  # per output, in physical units, on held-out samples
    loss: np.ndarray = None        # per training step: mean squared error of the normalised batch


def mesh_cell_values(mesh):
    """Cell values as the engine's Average_Point_Data_to_Cell forms them:
    the plain mean of each cell's unique vertices (the Delaunay generator's
    cells are simplices, so every vertex is unique)."""
    simp = mesh.simplices
    vel = mesh.vert_vel[simp].mean(axis=1)
    p = mesh.vert_p[simp].mean(axis=1)
    rho = mesh.vert_rho[simp].mean(axis=1)
    return np.concatenate([vel, p[:, None], rho[:, None]], axis=1)


def sample_mesh(mesh, n, seed=0, engine=None):
    """n uniform points in the mesh box and the values of their cells,
    located by the engine's mesh lookup when an engine holding the mesh is
    given, else by scipy."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    lo, hi = mesh.verts.min(axis=0), mesh.verts.max(axis=0)
    x = rng.uniform(lo, hi, size=(n, lo.size))
    if engine is not None:
        cell, vals = engine.mesh_eval(x)
        keep = cell >= 0
        return x[keep], vals[keep]
    if getattr(mesh, "tri", None) is not None:
        cell = mesh.tri.find_simplex(x)
        keep = cell >= 0
        return x[keep], mesh_cell_values(mesh)[cell[keep]]
    tri = Delaunay(mesh.verts)
    cell = tri.find_simplex(x)
    keep = cell >= 0
    # map the re-triangulation's simplices back to the mesh's cells
    key = {tuple(sorted(s)): k for k, s in enumerate(mesh.simplices.tolist())}
    remap = np.array([key[tuple(sorted(s))] for s in tri.simplices.tolist()])
    vals = mesh_cell_values(mesh)[remap[cell[keep]]]
    return x[keep], vals


def torch_forward(sg, x):
    """FP32 (torch, CPU) forward pass of a Surrogate at points x: the
    floating-point reference of the engine's kernel."""
    import torch

    dim = sg.in_shift.size
    w = torch.from_numpy(np.asarray(sg.weights, dtype=np.float32))
    b = torch.from_numpy(np.asarray(sg.biases, dtype=np.float32))
    h = torch.from_numpy(((x - sg.in_shift) / sg.in_scale).astype(np.float32))
    ow = ob = 0
    sizes = [(WIDTH, dim)] + [(WIDTH, WIDTH)] * (sg.n_hidden - 1)
    for (o, i) in sizes:
        W = w[ow:ow + o * i].reshape(o, i)
        h = torch.nn.functional.gelu(h @ W.T + b[ob:ob + o])
        ow += o * i
        ob += o
    W = w[ow:ow + (dim + 2) * WIDTH].reshape(dim + 2, WIDTH)
    y = h @ W.T + b[ob:ob + dim + 2]
    return y.double().numpy() * sg.out_scale + sg.out_shift


def init_params(dim, n_hidden, seed=0):
    """Initial weights and biases: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for
    every weight and bias of a layer (the usual default for dense layers)."""
    rng = np.random.default_rng(seed)
    sizes = [(WIDTH, dim)] + [(WIDTH, WIDTH)] * (n_hidden - 1) + [(dim + 2, WIDTH)]
    w = np.concatenate([rng.uniform(-1, 1, size=o * i) / np.sqrt(i) for o, i in sizes]).astype(np.float32)
    b = np.concatenate([rng.uniform(-1, 1, size=o) / np.sqrt(i) for o, i in sizes]).astype(np.float32)
    return w, b


def train(mesh, engine, n_samples=1_000_000, n_hidden=4, steps=3000, batch=16384, lr=3e-3, seed=0):
    """Fit the surrogate to the mesh field with the engine's own training
    kernels (fjsph_surrogate_train: BF16 MFMA forward and backward passes,
    FP32 Adam).  `engine` holds the mesh: it labels the samples with its
    mesh lookup and trains the network.  Returns a Surrogate whose
    train_rmse is the error on held-out samples (1 in 20), evaluated by the
    engine's inference kernel."""
    x, y = sample_mesh(mesh, n_samples, seed=seed, engine=engine)
    dim = x.shape[1]
    lo, hi = mesh.verts.min(axis=0), mesh.verts.max(axis=0)
    in_shift, in_scale = 0.5 * (lo + hi), 0.5 * (hi - lo)
    out_shift, out_scale = y.mean(axis=0), y.std(axis=0)
    out_scale = np.where(out_scale > 0, out_scale, 1.0)
    n_test = max(1, len(x) // 20)
    w0, b0 = init_params(dim, n_hidden, seed=seed)
    w, b, loss = engine.surrogate_train(x[n_test:], y[n_test:], n_hidden, w0, b0, in_shift, in_scale, out_shift,
                                        out_scale, steps=steps, batch=batch, lr=lr, seed=seed)
    sg = Surrogate(n_hidden, w, b, in_shift, in_scale, out_shift, out_scale)
    engine.set_surrogate(sg)
    err = engine.surrogate_eval(x[:n_test]) - y[:n_test]
    sg.train_rmse = np.sqrt(np.mean(err ** 2, axis=0))
    sg.loss = loss
    return sg


def random_surrogate(dim, n_hidden=4, seed=0, in_shift=None, in_scale=None):
    """A surrogate with random weights (for checks of the kernel against
    the FP32 forward)."""
    rng = np.random.default_rng(seed)
    sizes = [(WIDTH, dim)] + [(WIDTH, WIDTH)] * (n_hidden - 1) + [(dim + 2, WIDTH)]
    w = np.concatenate([rng.normal(0, 1.0 / np.sqrt(i), size=o * i) for o, i in sizes]).astype(np.float32)
    b = np.concatenate([rng.normal(0, 0.1, size=o) for o, _ in sizes]).astype(np.float32)
    return Surrogate(n_hidden, w, b,
                     np.zeros(dim) if in_shift is None else np.asarray(in_shift, dtype=np.float64),
                     np.ones(dim) if in_scale is None else np.asarray(in_scale, dtype=np.float64),
                     np.arange(dim + 2, dtype=np.float64), np.linspace(1.0, 2.0, dim + 2))
