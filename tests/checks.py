"""Comparison helpers shared by the oracle and GPU parity tests."""
import numpy as np


def rel_err(a, b):
    """max |a - b| / max |b| over one vector/matrix (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape or not np.all(np.isfinite(a)):
        return float("inf")  # a missing, reshaped or non-finite result never passes
    d = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / d) if d > 0 else float(np.max(np.abs(a - b), initial=0.0))


def max_rel_err(xs, ys, skip=()):
    """Largest per-step normwise relative error over paired lists."""
    worst = 0.0
    if len(xs) != len(ys):
        return float("inf")
    for i, (a, b) in enumerate(zip(xs, ys)):
        if i in skip or b is None:
            continue
        if a is None:
            return float("inf")
        e = rel_err(a, b)
        worst = e if not (e <= worst) else worst  # NaN propagates as a failure
    return worst


def elementwise_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
