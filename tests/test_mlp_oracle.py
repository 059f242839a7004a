"""The NumPy restatement of the surrogate's training pass (oracle/mlp_oracle.py)
pinned by central finite differences of the FP64 loss, and its BF16 mode
bounded against the FP64 gradient."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _net(dim, n_hidden, seed):
    rng = np.random.default_rng(seed)
    sizes = [(64, dim)] + [(64, 64)] * (n_hidden - 1) + [(dim + 2, 64)]
    w = np.concatenate([rng.normal(0, 1.0 / np.sqrt(i), size=o * i) for o, i in sizes])
    b = np.concatenate([rng.normal(0, 0.1, size=o) for o, _ in sizes])
    return w, b


@pytest.mark.parametrize("dim,n_hidden", [(2, 2), (3, 3)])
def test_gradient_matches_finite_differences(dim, n_hidden):
    rng = np.random.default_rng(dim * 10 + n_hidden)
    w, b = _net(dim, n_hidden, seed=dim)
    x = rng.uniform(-1, 1, size=(16, dim))
    t = rng.normal(size=(16, dim + 2))
    dw, db, loss = mo.forward_backward(w, b, n_hidden, dim, x, t, bf16=False)
    assert abs(loss - mo.loss_only(w, b, n_hidden, dim, x, t)) <= 1e-14 * abs(loss)
    h = 1e-6
    for arr, grad in ((w, dw), (b, db)):
        pick = rng.choice(arr.size, size=40, replace=False)
        for k in pick:
            ap, am = arr.copy(), arr.copy()
            ap[k] += h
            am[k] -= h
            wp, bp = (ap, b) if arr is w else (w, ap)
            wm, bm = (am, b) if arr is w else (w, am)
            fd = (mo.loss_only(wp, bp, n_hidden, dim, x, t) - mo.loss_only(wm, bm, n_hidden, dim, x, t)) / (2 * h)
            assert abs(fd - grad[k]) <= 1e-7 * max(1.0, np.abs(grad).max()), (k, fd, grad[k])


def test_gelu_grad_is_derivative():
    z = np.linspace(-6, 6, 1001)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(z))) < 1e-8


def test_bf16_round_known_values():
    # 1 + 2^-8 is halfway between BF16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265], dtype=np.float32)
    r = mo.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) == 0.0


def test_bf16_gradient_close_to_fp64():
    rng = np.random.default_rng(7)
    w, b = _net(3, 4, seed=3)
    x = rng.uniform(-1, 1, size=(256, 3))
    t = rng.normal(size=(256, 5))
    dw, db, _ = mo.forward_backward(w, b, 4, 3, x, t, bf16=False)
    dw16, db16, _ = mo.forward_backward(w, b, 4, 3, x, t, bf16=True)
    assert np.max(np.abs(dw16 - dw)) <= 3e-2 * np.max(np.abs(dw))
    assert np.max(np.abs(db16 - db)) <= 3e-2 * np.max(np.abs(db))
