"""fjsph_surrogate_grad_noalloc (device pointers, scratch in the library's
file-scope device array, no allocation) against fjsph_surrogate_grad (host
pointers, scratch allocated per call): the same kernels in the same order,
so the gradients and the loss must be bit-identical.  Two calls in a row on
one context, the second with a smaller batch (and a smaller network), so a
stale scratch word from the first call would show; then a batch above the
maximum, which must be refused without touching the outputs.
"""
import ctypes as C

import numpy as np
import pytest

from fjsph_amd import abi, surrogate

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """The HIP runtime the engine library is linked against (already loaded
    with it): device buffers from the same runtime, not from a second one
    that another package may bring into the process."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _chk(rc):
    assert rc == 0, f"HIP error {rc}"


def _dev(hip, a):
    a = np.ascontiguousarray(a)
    p = C.c_void_p()
    _chk(hip.hipMalloc(C.byref(p), max(a.nbytes, 1)))
    _chk(hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D))
    return p


def _host(hip, p, like):
    out = np.empty_like(like)
    _chk(hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H))
    return out


def _call_noalloc(eng, hip, n_hidden, w, b, xn, tn):
    ins = [_dev(hip, a) for a in (xn, tn, w, b)]
    nan_w = np.full(w.size, np.nan, np.float32)
    nan_b = np.full(b.size, np.nan, np.float32)
    nan_l = np.full(1, np.nan, np.float64)
    outs = [_dev(hip, a) for a in (nan_w, nan_b, nan_l)]
    _chk(hip.hipDeviceSynchronize())
    rc = eng.lib.fjsph_surrogate_grad_noalloc(eng.h, int(n_hidden), int(xn.shape[0]), *ins, *outs)
    _chk(hip.hipDeviceSynchronize())
    dw, db, loss = (_host(hip, p, a) for p, a in zip(outs, (nan_w, nan_b, nan_l)))
    for p in ins + outs:
        _chk(hip.hipFree(p))
    return rc, dw, db, float(loss[0])


@pytest.mark.parametrize("dim", [3, 2])
def test_noalloc_grad_matches_allocating_entry_point(dim):
    from fjsph_amd.engine import Engine
    s = abi.default_settings(dim)
    s.particle_step = 1e-4
    eng = Engine(s)
    hip = _hip()
    rng = np.random.default_rng(40 + dim)
    # (n_hidden, batch): the largest batch first, then a smaller call
    for n_hidden, batch in ((4, 4096), (2, 320)):
        w, b = surrogate.init_params(dim, n_hidden, seed=dim + n_hidden)
        b = (b + rng.normal(0, 0.2, size=b.size)).astype(np.float32)
        xn = rng.uniform(-1, 1, size=(batch, dim)).astype(np.float32)
        tn = rng.normal(size=(batch, dim + 2)).astype(np.float32)
        print(f"[noalloc grad] dim={dim} n_hidden={n_hidden} batch={batch}", flush=True)
        rw, rb, rloss = eng.surrogate_grad(n_hidden, w, b, xn, tn)
        rc, dw, db, loss = _call_noalloc(eng, hip, n_hidden, w, b, xn, tn)
        assert rc == abi.OK, eng.lib.fjsph_last_error(eng.h)
        assert np.array_equal(dw, rw), "weight gradients differ"
        assert np.array_equal(db, rb), "bias gradients differ"
        assert loss == rloss
        assert np.all(np.isfinite(dw)) and np.abs(dw).max() > 0
    # a batch above the maximum is refused before anything runs
    batch = 4096 + 64
    w, b = surrogate.init_params(dim, 2, seed=1)
    xn = np.zeros((batch, dim), np.float32)
    tn = np.zeros((batch, dim + 2), np.float32)
    rc, dw, db, loss = _call_noalloc(eng, hip, 2, w, b, xn, tn)
    assert rc == abi.ERR_ARG
    assert np.all(np.isnan(dw)) and np.isnan(loss)
    eng.close()
