"""Comparison helpers shared by the oracle and GPU parity tests."""
import numpy as np


def rel_err(a, b):
    """max |a - b| / max |b| over one vector/matrix (normwise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / d) if d > 0 else float(np.max(np.abs(a - b), initial=0.0))


def max_rel_err(xs, ys, skip=()):
    """Largest per-step normwise relative error over paired lists."""
    worst = 0.0
    for i, (a, b) in enumerate(zip(xs, ys)):
        if i in skip or a is None or b is None:
            continue
        worst = max(worst, rel_err(a, b))
    return worst


def elementwise_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
