"""Comparison helpers for parity tests."""
import numpy as np


def rel_err(a, b) -> float:
    """max |a - b| / max |b| (field-wise relative error; 0 when both are zero)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    d = float(np.max(np.abs(a - b))) if b.size else 0.0
    if scale == 0.0:
        return d
    return d / scale


def assert_close(a, b, tol, what=""):
    e = rel_err(a, b)
    assert e <= tol, f"{what}: relative error {e:.3e} > {tol:.1e}"
    return e
