"""The MLP surrogate kernel (BF16 MFMA, FP32 accumulation) against a torch
reference of the same network, and the surrogate aero source in a step.

Tolerances (outputs normalised by their scale):
  vs the BF16-emulating reference (weights and layer inputs rounded to BF16,
  FP32 sums): median 1e-4, max 3e-2 -- the sums' order differs, which can
  move a BF16 rounding of an activation by one unit;
  vs plain FP32 torch: max 5e-2 (BF16 operands);
  surrogate aero field vs the mesh lookup it was trained on: 5% of the
  stream speed (RMS), the network being a smooth fit of a piecewise-constant
  field.

The training kernels (fjsph_surrogate_grad) against the NumPy oracle of
oracle/mlp_oracle.py (itself pinned by finite differences): with BF16
rounding emulated at the kernels' rounding points, normwise 2e-2 (FP32 sums
in another order can move a BF16 rounding of an activation or of dZ by one
unit); against the exact FP64 gradient, normwise 5e-2.
"""
import numpy as np
import pytest

from fjsph_amd import abi, cases, surrogate

pytestmark = pytest.mark.gpu


def _bf16(a):
    import torch
    return torch.as_tensor(a, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def _emulated_forward(sg, x):
    import torch
    dim = sg.in_shift.size
    w = torch.from_numpy(np.asarray(sg.weights, dtype=np.float32))
    b = torch.from_numpy(np.asarray(sg.biases, dtype=np.float32))
    h = _bf16(((x - sg.in_shift) / sg.in_scale).astype(np.float32))
    ow = ob = 0
    for (o, i) in [(64, dim)] + [(64, 64)] * (sg.n_hidden - 1):
        W = _bf16(w[ow:ow + o * i].reshape(o, i))
        h = _bf16(torch.nn.functional.gelu(h @ W.T + b[ob:ob + o]))
        ow += o * i
        ob += o
    W = _bf16(w[ow:ow + (dim + 2) * 64].reshape(dim + 2, 64))
    return (h @ W.T + b[ob:ob + dim + 2]).double().numpy()


@pytest.mark.parametrize("dim,n_hidden", [(3, 4), (2, 3)])
def test_surrogate_kernel_matches_torch(dim, n_hidden):
    from fjsph_amd.engine import Engine
    s = abi.default_settings(dim)
    s.particle_step = 1e-4
    eng = Engine(s)
    sg = surrogate.random_surrogate(dim, n_hidden, seed=dim, in_shift=np.full(dim, 0.5), in_scale=np.full(dim, 0.5))
    eng.set_surrogate(sg)
    x = np.random.default_rng(7).uniform(0.0, 1.0, size=(5000, dim))
    got = (eng.surrogate_eval(x) - sg.out_shift) / sg.out_scale
    emu = _emulated_forward(sg, x)
    ref = (surrogate.torch_forward(sg, x) - sg.out_shift) / sg.out_scale
    scale = np.abs(ref).max()
    d_emu = np.abs(got - emu) / scale
    d_ref = np.abs(got - ref) / scale
    print(f"dim {dim}: vs emulated median {np.median(d_emu):.2e} max {d_emu.max():.2e}; vs fp32 max {d_ref.max():.2e}",
          flush=True)
    assert np.median(d_emu) <= 1e-4 and d_emu.max() <= 3e-2
    assert d_ref.max() <= 5e-2
    eng.close()


def test_surrogate_aero_step():
    """A droplet with the surrogate aero source, trained on its mesh: the
    aero particles' gas velocity tracks the mesh lookup, and steps run."""
    from fjsph_amd.engine import Engine
    base = cases.with_mesh(cases.droplet_3d(diameter=1.0e-3, dx=5e-5), n_points=20000)
    lab = Engine(base.settings)
    lab.set_mesh(base.mesh)
    sg = surrogate.train(base.mesh, lab, n_samples=200_000, steps=1500, batch=8192)
    lab.close()
    print("surrogate held-out rmse:", sg.train_rmse, "loss", sg.loss[0], "->", sg.loss[-1], flush=True)
    assert sg.loss[-100:].mean() < 0.1 * sg.loss[:10].mean()
    eng_m = Engine(base)
    eng_m.build_neighbours()
    eng_m.dsph_prestep()
    eng_m.aero_velocity()
    vm, cid = eng_m.field("CELLV"), eng_m.field("CELLID")
    base.settings.aero_source = abi.SURROGATE
    eng_s = Engine(base)
    eng_s.set_surrogate(sg)
    eng_s.build_neighbours()
    eng_s.dsph_prestep()
    eng_s.aero_velocity()
    vs, cis = eng_s.field("CELLV"), eng_s.field("CELLID")
    aero = cid >= 0
    assert aero.sum() > 100 and np.array_equal(aero, cis >= 0)
    speed = float(base.settings.v_inf[0])
    rms = np.sqrt(np.mean(np.sum((vs[aero] - vm[aero]) ** 2, axis=1)))
    print(f"surrogate vs mesh gas velocity rms {rms:.3f} m/s of {speed} m/s", flush=True)
    assert rms < 0.05 * speed
    eng_s.integrate_no_update()
    for _ in range(2):
        info = eng_s.integrate_step()
        assert np.isfinite(info["rms_error"])
    assert np.all(np.isfinite(eng_s.field("X")))


def test_mesh_eval_matches_scipy():
    """fjsph_mesh_eval locates points in the same simplices as scipy's
    Delaunay.find_simplex and returns those cells' averaged values."""
    from scipy.spatial import Delaunay

    from fjsph_amd.engine import Engine
    s = abi.default_settings(3)
    s.particle_step = 1e-4
    eng = Engine(s)
    mesh = cases.delaunay_mesh([0, 0, 0], [1e-2, 2e-3, 2e-3], 20000, seed=3)
    eng.set_mesh(mesh)
    x = np.random.default_rng(4).uniform([0, 0, 0], [1e-2, 2e-3, 2e-3], size=(20000, 3))
    cell, vals = eng.mesh_eval(x)
    ref = mesh.tri.find_simplex(x)
    assert np.array_equal(cell >= 0, ref >= 0)
    ok = cell >= 0
    assert ok.sum() > 19000
    assert np.array_equal(cell[ok], ref[ok])
    _, cvel, cp, crho = eng.mesh_cells()
    assert np.array_equal(vals[ok, :3], cvel[cell[ok]])
    assert np.array_equal(vals[ok, 3], cp[cell[ok]])
    del Delaunay
    eng.close()


def _normwise(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("dim,n_hidden", [(3, 4), (2, 2)])
def test_training_gradients_match_oracle(dim, n_hidden):
    """One forward/backward pass of the training kernels against the oracle."""
    from fjsph_amd.engine import Engine
    from oracle import mlp_oracle as mo
    s = abi.default_settings(dim)
    s.particle_step = 1e-4
    eng = Engine(s)
    rng = np.random.default_rng(11 * dim + n_hidden)
    w, b = surrogate.init_params(dim, n_hidden, seed=dim)
    b = (b + rng.normal(0, 0.2, size=b.size)).astype(np.float32)
    batch = 256
    xn = rng.uniform(-1, 1, size=(batch, dim)).astype(np.float32)
    tn = rng.normal(size=(batch, dim + 2)).astype(np.float32)
    print(f"[train grad] dim={dim} n_hidden={n_hidden} batch={batch}", flush=True)
    dw, db, loss = eng.surrogate_grad(n_hidden, w, b, xn, tn)
    eng.close()
    rw, rb, rloss = mo.forward_backward(w, b, n_hidden, dim, xn.astype(np.float64), tn.astype(np.float64))
    fw, fb, floss = mo.forward_backward(w, b, n_hidden, dim, xn.astype(np.float64), tn.astype(np.float64),
                                        bf16=False)
    errs = {"dW_bf16": _normwise(dw, rw), "db_bf16": _normwise(db, rb),
            "dW_fp64": _normwise(dw, fw), "db_fp64": _normwise(db, fb)}
    print("gradient errors:", {k: f"{v:.1e}" for k, v in errs.items()}, "loss", loss, rloss, floss, flush=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    assert errs["dW_bf16"] <= 2e-2 and errs["db_bf16"] <= 2e-2
    assert errs["dW_fp64"] <= 5e-2 and errs["db_fp64"] <= 5e-2


def test_training_reduces_loss_and_fits_mesh():
    """fjsph_surrogate_train on a mesh field: the loss falls, and the trained
    network's gas velocity on held-out samples is within 1% of the stream
    speed (RMS) of the mesh lookup's."""
    from fjsph_amd.engine import Engine
    base = cases.with_mesh(cases.droplet_3d(diameter=1.0e-3, dx=5e-5), n_points=20000)
    lab = Engine(base.settings)
    lab.set_mesh(base.mesh)
    sg = surrogate.train(base.mesh, lab, n_samples=400_000, steps=8000, batch=16384)
    lab.close()
    speed = float(base.settings.v_inf[0])
    vel_rms = float(np.sqrt(np.sum(sg.train_rmse[:3] ** 2)))
    print(f"loss {sg.loss[0]:.3e} -> {sg.loss[-1]:.3e}; held-out gas velocity rms {vel_rms:.4f} m/s "
          f"({100 * vel_rms / speed:.2f}% of {speed} m/s)", flush=True)
    assert sg.loss[-100:].mean() < 0.05 * sg.loss[:10].mean()
    assert vel_rms < 0.01 * speed


def test_mesh_eval_structured_matches_closed_form():
    """fjsph_mesh_eval on the structured tetrahedral mesh returns the cell
    of the closed form (cube, then axis order of the fractional
    coordinates) for every point."""
    from fjsph_amd.engine import Engine
    from tests.test_oracle import _kuhn_cell
    s = abi.default_settings(3)
    s.particle_step = 1e-4
    eng = Engine(s)
    lo, hi = np.zeros(3), np.array([1e-2, 4e-3, 3e-3])
    mesh = cases.kuhn_mesh(lo, hi, 200_000)
    eng.set_mesh(mesh)
    n = np.array([len(np.unique(mesh.verts[:, k])) - 1 for k in range(3)])
    x = np.random.default_rng(6).uniform(lo, hi, size=(50_000, 3))
    cell, vals = eng.mesh_eval(x)
    ref = _kuhn_cell(lo, (hi - lo) / n, n, x)
    print(f"[mesh eval structured] cells {mesh.ncells}, mismatches {(cell != ref).sum()}", flush=True)
    assert np.array_equal(cell, ref)
    _, cvel, cp, _ = eng.mesh_cells()
    assert np.array_equal(vals[:, :3], cvel[cell])
    eng.close()


def test_surrogate_lookup_gain_over_mesh():
    """What the trained network gains over the method without it, on the
    4M-particle coflow jet with its 20M-cell gas mesh: within whole steps
    (the particles move between lookups), the surrogate gives the aero
    particles' gas state in less time than the mesh lookup (FindCell), at
    under 1% gas-velocity error (RMS, of the stream speed) on the same
    particles, with 26 KB of weights in place of the mesh; engine aero
    timers over 3 steps each.  (Repeated lookups at fixed positions hit
    every particle's cached cell, and there the mesh lookup is the faster
    one.)"""
    from fjsph_amd.engine import Engine
    case = cases.with_mesh(cases.jet_column_3d(n_target=4_000_000), kind="kuhn", n_cells=20_000_000)
    print(f"[gain] jet n={case.n}, mesh cells {case.mesh.ncells}", flush=True)
    lab = Engine(case.settings)
    lab.set_mesh(case.mesh)
    sg = surrogate.train(case.mesh, lab, n_samples=1_000_000, steps=3000, batch=16384)
    lab.close()
    out = {}
    for name in ("mesh", "surrogate"):
        if name == "surrogate":
            case.settings.aero_source = abi.SURROGATE
        eng = Engine(case)
        if name == "surrogate":
            eng.set_surrogate(sg)
        eng.build_neighbours()
        eng.dsph_prestep()
        eng.aero_velocity()
        fields = (eng.field("CELLV"), eng.field("CELLID"))
        eng.integrate_no_update()
        eng.integrate_step()
        eng.enable_timers(True)
        for _ in range(3):
            eng.integrate_step()
        ms, calls = eng.timers()["aero"]
        out[name] = (ms / calls,) + fields
        print(f"[gain] {name}: {calls} lookups, {ms / calls * 1e3:.1f} us each", flush=True)
        eng.close()
    aero = out["mesh"][2] >= 0
    assert aero.sum() > 100_000 and np.array_equal(aero, out["surrogate"][2] >= 0)
    speed = float(case.settings.v_inf[0])
    d = out["surrogate"][1][aero] - out["mesh"][1][aero]
    rms = float(np.sqrt(np.mean(np.sum(d ** 2, axis=1))))
    t_mesh, t_sg = out["mesh"][0], out["surrogate"][0]
    print(f"aero particles {aero.sum()}: mesh lookup {t_mesh * 1e3:.1f} us, surrogate {t_sg * 1e3:.1f} us "
          f"({t_mesh / t_sg:.2f}x); gas velocity rms {rms:.4f} m/s ({100 * rms / speed:.2f}%)", flush=True)
    assert rms < 0.01 * speed
    assert t_sg < t_mesh
