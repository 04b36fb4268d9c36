"""Covariance output formats (ukhip_batch_download_cov, UKHIP_COVARIANCE_W):
the diagonal 'D' and the reference's 'W' representation, W = L^{-1} for the
Cholesky factor S = L L^T (c/kalman_base.c:61-175: W^T W = C^{-1}).

Bars: 'D' equals the diagonal of the full covariance exactly (a gather);
'W' is lower triangular and W S W^T = I to 1e-10; through the C ABI with the
option bit, covariance_type is 'W' and W S_oracle W^T = I to 1e-9.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from ukhip import capi
from ukhip import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    device.lib()
    return device


@pytest.mark.parametrize("n,m", [(6, 6), (40, 20)])
def test_batch_formats(dev, n, m):
    F = 0.98 * P.random_orthogonal(P.rng_for(700 + n), n)
    # k = 1 only where one step determines the state (m = n): with m < n a
    # single step leaves S singular, and W does not exist
    lengths = (1, 5, 257) if m == n else (5, 257)
    probs = [P.iid_problem(k, n, m, seed=710 + k, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name=f"n{n}_k{k}") for k in lengths]
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        for i, pr in enumerate(probs):
            _, S = b.download(i)
            D = b.download_cov(i, "D")
            assert np.array_equal(D, np.stack([np.diag(s) for s in S]))
            Wt = b.download_cov(i, "W")
            worst = 0.0
            for s, W in zip(S, Wt):
                assert np.all(np.isfinite(W)), pr.name
                assert np.array_equal(W, np.tril(W))
                worst = max(worst, float(np.max(np.abs(W @ s @ W.T - np.eye(n)))))
            say(f"  {pr.name}: max |W S W^T - I| = {worst:.2e}")
            assert worst < 1e-10, pr.name
    finally:
        b.free()


def test_c_abi_w_representation(dev, oracle):
    pr = P.random_problem(91, 30, 5, 3, K_kind="W", C_kind="C", unobserved=(4, 7))
    L = dev.lib()
    eqs = capi.EquationSet(pr)
    L.ukhip_clear_error()
    L.kalman_smooth_oddeven(capi.UKHIP_COVARIANCE_W, eqs.ptrs, pr.k)
    assert L.ukhip_last_error() == b""
    types = [eqs.structs[i].covariance_type for i in range(pr.k)]
    assert all(t == b"W" for t in types), types
    _, Ws = eqs.collect(L.matrix_free)
    _, So = oracle.oddeven(pr)
    worst = max(float(np.max(np.abs(W @ s @ W.T - np.eye(pr.n[i])))) for i, (W, s) in enumerate(zip(Ws, So)))
    say(f"C ABI 'W' covariances: max |W S_oracle W^T - I| = {worst:.2e}")
    assert worst < 1e-9


UKHIP_COVARIANCE_DIAG = 1 << 25


@pytest.mark.parametrize("n,m", [(6, 6), (13, 5), (40, 20), (64, 32)])
def test_diagonal_only_path(dev, oracle, n, m):
    """UKHIP_COVARIANCE_DIAG: the back substitution of the first level forms
    only diag(S_j) (vector units for n <= 16, matrix cores above). The
    diagonals equal the full path's to rounding (1e-13) and the oracle's to
    1e-9; the estimates are bit-identical to the full path's; odd and even
    lengths and a one-step problem."""
    F = 0.98 * P.random_orthogonal(P.rng_for(740 + n), n)
    lengths = (1, 6, 257, 1024) if m == n else (6, 257, 1024)
    probs = [P.iid_problem(k, n, m, seed=750 + k, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name=f"n{n}_k{k}") for k in lengths]
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        full = [(b.download(i)[0], b.download_cov(i, "D")) for i in range(len(probs))]
        b.smooth("oddeven", options=UKHIP_COVARIANCE_DIAG)
        for i, pr in enumerate(probs):
            x = b.download(i, want_cov=False)[0]
            D = b.download_cov(i, "D")
            assert np.array_equal(x, full[i][0]), pr.name
            d_full = float(np.max(np.abs(D - full[i][1]) / np.abs(full[i][1])))
            _, oc = oracle.oddeven(pr)
            d_or = float(np.max(np.abs(D - np.stack([np.diag(c) for c in oc])) / np.abs(full[i][1])))
            say(f"  {pr.name}: diagonals vs full path {d_full:.1e}, vs oracle {d_or:.1e}")
            assert d_full < 1e-13 and d_or < 1e-9, pr.name
    finally:
        b.free()
