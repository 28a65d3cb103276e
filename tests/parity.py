"""Comparison helpers for GPU-versus-oracle parity."""
import numpy as np


def compare(gpu, ref, rtol=1e-12):
    """Return (max scaled error, fraction bit-identical).

    Scaled error of an entry = |g - r| / max(|r|, 1e-3 * max|r_row|): plain
    relative error, except that entries which cancel to near zero are
    measured against their field's magnitude."""
    g = np.atleast_2d(np.asarray(gpu, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    assert g.shape == r.shape, (g.shape, r.shape)
    rowmax = np.abs(r).max(axis=-1, keepdims=True)
    scale = np.maximum(np.abs(r), 1e-3 * rowmax)
    scale = np.where(scale > 0, scale, 1.0)
    err = np.abs(g - r) / scale
    err = np.where(np.isnan(g) & np.isnan(r), 0.0, err)
    same = (g.view(np.int64) == r.view(np.int64)).mean()
    return float(np.nanmax(err)) if err.size else 0.0, float(same)


def assert_parity(gpu, ref, rtol=1e-12, what=""):
    err, same = compare(gpu, ref, rtol)
    assert err <= rtol, f"{what}: max scaled error {err:.3e} > {rtol:g} (bit-identical {same:.6f})"
    return err, same
