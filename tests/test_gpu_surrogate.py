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
    sg = surrogate.train(base.mesh, n_samples=200_000, steps=1500, batch=8192)
    print("surrogate held-out rmse:", sg.train_rmse, flush=True)
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
