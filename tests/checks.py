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


def assoc_parity(got, oracle_assoc, exact, tol=1e-9, factor=10.0, floor=1e-12):
    """Parity of an associative-scan result with the oracle's.

    The associative smoother works in covariance form; where the filtered
    covariances are ill-conditioned (an information-form prior inverted when
    m < n, slowly decaying dynamics) the reference algorithm itself departs
    from the least-squares solution by up to ~1e-7 (the oracle and the
    reference fixtures agree on that). Two implementations of it then differ
    by rounding amplified the same way. So: within `tol` of the oracle, or, if
    not, no further from the exact least-squares result (`exact`, QR-based
    odd-even) than `factor` times the oracle's own distance plus `floor`.

    Each argument is a (states, covariances) pair of per-step lists.
    Returns (ok, description)."""
    d = max(max_rel_err(got[0], oracle_assoc[0]), max_rel_err(got[1], oracle_assoc[1]))
    if d < tol:
        return True, f"vs oracle {d:.2e}"
    eg = max(max_rel_err(got[0], exact[0]), max_rel_err(got[1], exact[1]))
    eo = max(max_rel_err(oracle_assoc[0], exact[0]), max_rel_err(oracle_assoc[1], exact[1]))
    return eg <= factor * eo + floor, f"vs oracle {d:.2e}; from least squares: {eg:.2e} (oracle's own {eo:.2e})"
