"""Shared comparison helpers for the engine-vs-oracle parity tests."""
import numpy as np


def normwise_err(gpu, ref):
    """max_i |g_i - r_i| / max_i |r_i| (0 when both are zero)."""
    g = np.asarray(gpu, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    scale = np.max(np.abs(r)) if r.size else 0.0
    diff = np.max(np.abs(g - r)) if r.size else 0.0
    if scale == 0.0:
        return float(diff)
    return float(diff / scale)


def assert_close(name, gpu, ref, rtol, sel=None):
    g = np.asarray(gpu)
    r = np.asarray(ref)
    if sel is not None:
        g, r = g[sel], r[sel]
    assert np.all(np.isfinite(g)), f"{name}: non-finite engine values"
    e = normwise_err(g, r)
    assert e <= rtol, f"{name}: normwise relative error {e:.3e} > {rtol:.1e}"
    return e


def same_neighbour_sets(off_a, idx_a, off_b, idx_b):
    if not np.array_equal(off_a, off_b):
        return False
    return np.array_equal(idx_a, idx_b)
