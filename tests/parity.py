"""Comparison helpers for GPU-versus-oracle parity."""
import numpy as np


def compare(gpu, ref, rtol=1e-12):
    """Return (max scaled error, fraction bit-identical).

    Scaled error of an entry = |g - r| / max(|r|, 1e-3 * max|r_row|): plain
    relative error, except that entries which cancel to near zero are
    measured against their field's magnitude.  Non-finite entries must match
    exactly: NaN against NaN and an infinity against the same infinity count
    as equal; NaN or an infinity against anything else is an infinite
    error."""
    g = np.atleast_2d(np.asarray(gpu, dtype=np.float64))
    r = np.atleast_2d(np.asarray(ref, dtype=np.float64))
    assert g.shape == r.shape, (g.shape, r.shape)
    if g.size == 0:
        return 0.0, 1.0
    fin_r = np.where(np.isfinite(r), r, 0.0)
    rowmax = np.abs(fin_r).max(axis=-1, keepdims=True)
    scale = np.maximum(np.abs(fin_r), 1e-3 * rowmax)
    scale = np.where(scale > 0, scale, 1.0)
    finite = np.isfinite(g) & np.isfinite(r)
    with np.errstate(invalid="ignore", over="ignore"):
        err = np.where(finite, np.abs(g - r) / scale, np.inf)
    same_nonfinite = (np.isnan(g) & np.isnan(r)) | (np.isinf(g) & (g == r))
    err = np.where(same_nonfinite, 0.0, err)
    same = (g.view(np.int64) == r.view(np.int64)).mean()
    return float(np.max(err)), float(same)


def assert_parity(gpu, ref, rtol=1e-12, what=""):
    err, same = compare(gpu, ref, rtol)
    assert err <= rtol, f"{what}: max scaled error {err:.3e} > {rtol:g} (bit-identical {same:.6f})"
    return err, same
