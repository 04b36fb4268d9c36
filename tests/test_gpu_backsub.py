"""The covariance back substitution on the FP64 matrix cores
(k_backsub_mfma, n > 16) against the vector-unit kernel (k_backsub) and the
oracle, for state dimensions at, below and between the template buckets
(padding inside a bucket), with unobserved steps, m < n and odd lengths.

Bars: each path within 1e-9 (per-step normwise relative error) of the oracle
on estimates and covariances; the two paths within 1e-11 of each other (the
same quantities in another summation order: W^T = Sigma T^T, explicit
R^{-1} products instead of substitutions).
"""
import os
import sys

import numpy as np
import pytest

from checks import max_rel_err
from ukhip import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    device.lib()
    return device


def _both(dev, probs):
    L = dev.lib()
    out = {}
    try:
        L.ukhip_use_graphs(0)
        for flag in (1, 0):
            L.ukhip_use_mfma(flag)
            b = dev.Batch(probs)
            b.smooth("oddeven")
            out[flag] = [b.download(i) for i in range(len(probs))]
            b.free()
    finally:
        L.ukhip_use_mfma(1)
        L.ukhip_use_graphs(1)
    return out


@pytest.mark.parametrize("n,m", [(17, 5), (24, 12), (32, 32), (40, 20), (48, 24), (57, 30), (64, 32)])
def test_mfma_matches_vector_and_oracle(dev, oracle, n, m):
    F = 0.98 * P.random_orthogonal(P.rng_for(300 + n), n)
    probs = [P.iid_problem(k, n, m, seed=400 + n + k, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name=f"n{n}m{m}_k{k}")
             for k in (1, 2, 7, 130, 333)]
    say(f"n={n} m={m}: {len(probs)} problems")
    out = _both(dev, probs)
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    try:
        for i, pr in enumerate(probs):
            (xm, cm), (xv, cv) = out[1][i], out[0][i]
            d = max(max_rel_err(list(xm), list(xv)), max_rel_err(list(cm), list(cv)))
            ox, oc = oracle.oddeven(pr)
            em = max(max_rel_err(list(xm), ox), max_rel_err(list(cm), oc))
            ev = max(max_rel_err(list(xv), ox), max_rel_err(list(cv), oc))
            say(f"  {pr.name}: matrix cores vs oracle {em:.2e}, vector {ev:.2e}, between {d:.2e}")
            assert em < TOL and ev < TOL and d < 1e-11, pr.name
    finally:
        oracle.lib.oracle_set_threads(1)


def test_ill_conditioned_n64(dev, oracle):
    """The precision workload's dynamics (F = Q diag(logspace(0,-3,64)),
    K = 1e-2 I) at a length the oracle checks in seconds."""
    F = P.random_orthogonal(P.rng_for(1004), 64) @ np.diag(np.logspace(0, -3, 64))
    pr = P.iid_problem(515, 64, 32, seed=14, F=F, G_scale=1 / 8.0, Kvar=1e-2, Cvar=1.0, name="n64_illcond_k515")
    out = _both(dev, [pr])
    ox, oc = oracle.oddeven(pr)
    (xm, cm), (xv, cv) = out[1][0], out[0][0]
    em = max(max_rel_err(list(xm), ox), max_rel_err(list(cm), oc))
    ev = max(max_rel_err(list(xv), ox), max_rel_err(list(cv), oc))
    say(f"ill-conditioned n=64: matrix cores {em:.2e}, vector {ev:.2e}")
    assert em < TOL and ev < TOL


@pytest.mark.parametrize("n,m", [(17, 5), (40, 20), (64, 32)])
def test_estimates_only_path(dev, oracle, n, m):
    """KALMAN_NO_COVARIANCE (one wave per step, k_estimate) gives the
    estimates of the covariance path (another summation order: within 1e-12)
    and of the oracle (1e-9)."""
    from ukhip import capi
    F = 0.97 * P.random_orthogonal(P.rng_for(500 + n), n)
    probs = [P.iid_problem(k, n, m, seed=600 + n + k, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name=f"n{n}m{m}_k{k}") for k in (1, 2, 9, 301)]
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        full = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
        b.smooth("oddeven", options=capi.KALMAN_NO_COVARIANCE)
        est = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
    finally:
        b.free()
    for i, pr in enumerate(probs):
        d = max_rel_err(list(est[i]), list(full[i]))
        ox, _ = oracle.oddeven(pr, capi.KALMAN_NO_COVARIANCE)
        e = max_rel_err(list(est[i]), ox)
        say(f"  {pr.name}: estimates-only vs covariance path {d:.2e}, vs oracle {e:.2e}")
        assert d < 1e-12 and e < TOL, pr.name
