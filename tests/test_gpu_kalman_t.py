"""kalman_t on the ingest path: kalman_evolve / kalman_observe write each
step into the device layout's fields, kalman_smooth uploads them as they are.

Covers what the ingest changes: representations that widen mid-sequence
('w' then full K and C, more observation rows, a non-identity H appearing),
kalman_forget (the earliest remaining step loses its evolution equation),
kalman_rollback (a step reopened and observed anew), and the 'W' covariance
option. Bar: every estimate and covariance within 1e-9 of the oracle on the
equivalent explicit problem.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from checks import max_rel_err, rel_err
from ukhip import capi
from ukhip import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    from ukhip import device
    L = device.lib()
    L.kalman_create_options.restype = C.c_void_p
    L.kalman_create_options.argtypes = [C.c_uint32]
    L.kalman_evolve.argtypes = [C.c_void_p, C.c_int32] + [C.POINTER(capi.KalmanMatrix)] * 4 + [C.c_char]
    L.kalman_observe.argtypes = [C.c_void_p] + [C.POINTER(capi.KalmanMatrix)] * 3 + [C.c_char]
    L.kalman_smooth.argtypes = [C.c_void_p]
    L.kalman_estimate.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_estimate.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_covariance.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_covariance.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_covariance_type.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_covariance_type.restype = C.c_char
    L.kalman_forget.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_rollback.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_earliest.argtypes = [C.c_void_p]
    L.kalman_earliest.restype = C.c_int32
    L.kalman_latest.argtypes = [C.c_void_p]
    L.kalman_latest.restype = C.c_int32
    L.kalman_free.argtypes = [C.c_void_p]
    return L


def _concat(a, b):
    """Steps of a, then steps of b; b's first step gets the evolution equation
    of b's second (it follows a's last step)."""
    f = {x: list(getattr(a, x)) + list(getattr(b, x))
         for x in ("n", "H", "F", "c", "K", "K_type", "G", "o", "C", "C_type")}
    for x in ("H", "F", "c", "K", "K_type"):
        f[x][a.k] = getattr(b, x)[1]
    return P.Problem(k=a.k + b.k, name=f"{a.name}+{b.name}", **f)


def _sub(pr, lo, hi):
    """Steps lo..hi-1 as a problem of their own: step lo loses its evolution equation."""
    f = {x: getattr(pr, x)[lo:hi] for x in ("n", "H", "F", "c", "K", "K_type", "G", "o", "C", "C_type")}
    for x in ("H", "F", "c", "K", "K_type"):
        f[x] = [None] + f[x][1:]
    return P.Problem(k=hi - lo, name=f"{pr.name}[{lo}:{hi}]", **f)


def _feed(L, kal, eqs, i):
    s = eqs.structs[i]
    L.kalman_evolve(kal, eqs.problem.n[i], s.H, s.F, s.c, s.K, s.K_type or b"W")
    L.kalman_observe(kal, s.G, s.o, s.C, s.C_type or b"W")


def _check(L, kal, pr, first, oracle, what):
    ox, oc = oracle.oddeven(pr)
    worst = 0.0
    for i in range(pr.k):
        e = L.kalman_estimate(kal, first + i)
        cm = L.kalman_covariance(kal, first + i)
        x, cv = capi._read(e)[:, 0], capi._read(cm)
        L.matrix_free(e)
        L.matrix_free(cm)
        worst = max(worst, rel_err(x, ox[i]), rel_err(cv, oc[i]))
    say(f"  {what}: worst {worst:.2e}")
    assert worst < TOL, what


def test_widening_forget_rollback(L, oracle):
    n = 5
    a = P.random_problem(301, 12, n, 2, K_kind="w", C_kind="w", unobserved=(3,))
    b = P.random_problem(302, 14, n, 4, K_kind="C", C_kind="W", H_kind="R", unobserved=(6,))
    full = _concat(a, b)
    eqs = capi.EquationSet(full)
    kal = L.kalman_create_options(capi.KALMAN_ALGORITHM_ODDEVEN)
    try:
        for i in range(full.k):
            _feed(L, kal, eqs, i)
        assert (L.kalman_earliest(kal), L.kalman_latest(kal)) == (0, full.k - 1)
        e = L.kalman_estimate(kal, 3)
        assert np.all(np.isnan(capi._read(e)))  # not smoothed yet
        L.matrix_free(e)
        L.kalman_smooth(kal)
        _check(L, kal, full, 0, oracle, "widened representations (w -> C/W, m 2 -> 4, H = I -> random)")

        L.kalman_forget(kal, 4)
        assert L.kalman_earliest(kal) == 5
        L.kalman_smooth(kal)
        _check(L, kal, _sub(full, 5, full.k), 5, oracle, "after kalman_forget(4)")

        L.kalman_rollback(kal, 15)
        assert L.kalman_latest(kal) == 14
        other = P.random_problem(303, full.k, n, 4, K_kind="C", C_kind="W", H_kind="R")
        eq_other = capi.EquationSet(other)  # owns the buffers the structs point into
        s_other = eq_other.structs
        L.kalman_observe(kal, s_other[15].G, s_other[15].o, s_other[15].C, b"W")  # step 15 observed anew
        assert L.kalman_latest(kal) == 15
        L.kalman_smooth(kal)
        redo = _sub(full, 5, 16)
        redo.G[-1], redo.o[-1], redo.C[-1], redo.C_type[-1] = other.G[15], other.o[15], other.C[15], "W"
        _check(L, kal, redo, 5, oracle, "after kalman_rollback(15) and a new observation")
    finally:
        L.kalman_free(kal)


@pytest.mark.parametrize("algo", ["oddeven", "associative"])
def test_w_covariance_option(L, oracle, algo):
    pr = P.random_problem(311, 23, 4, 4, K_kind="W", C_kind="C")
    eqs = capi.EquationSet(pr)
    flag = capi.KALMAN_ALGORITHM_ODDEVEN if algo == "oddeven" else capi.KALMAN_ALGORITHM_ASSOCIATIVE
    kal = L.kalman_create_options(flag | capi.UKHIP_COVARIANCE_W)
    try:
        for i in range(pr.k):
            _feed(L, kal, eqs, i)
        L.kalman_smooth(kal)
        _, oc = oracle.oddeven(pr)
        worst = 0.0
        for i in range(pr.k):
            assert L.kalman_covariance_type(kal, i) == b"W"
            cm = L.kalman_covariance(kal, i)
            W = capi._read(cm)
            L.matrix_free(cm)
            worst = max(worst, float(np.max(np.abs(W @ oc[i] @ W.T - np.eye(4)))))
        say(f"  {algo}: max |W S_oracle W^T - I| = {worst:.2e}")
        assert worst < 1e-9
    finally:
        L.kalman_free(kal)


def test_shapes_through_kalman_t(L, oracle):
    """kalman_evolve / kalman_observe with the reference's per-step shapes
    (kalmanTest.m:80-95): a regular stretch first, then a change of state
    dimension, rectangular H and more observation rows than states switch the
    ingest to its general form; smoothing, kalman_forget and kalman_rollback
    then work on it. Estimates and covariances (of each step's own dimension)
    against the oracle."""
    import json
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shapes_oddeven.npz")
    d = np.load(gold)
    meta = json.loads(bytes(d["meta"]).decode())
    for tag in ("grow8_wC", "rand2_k64_wC", "wide16_wC", "mid3obs_UW"):
        assert tag in meta["cases"]
        pre = tag + "__"
        pr = P.problem_from_arrays({k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}, tag)
        eqs = capi.EquationSet(pr)
        kal = L.kalman_create_options(capi.KALMAN_ALGORITHM_ODDEVEN)
        try:
            for i in range(pr.k):
                _feed(L, kal, eqs, i)
            L.kalman_smooth(kal)
            _check(L, kal, pr, 0, oracle, f"{tag}: all steps")
            if pr.k >= 8:
                L.kalman_forget(kal, 2)
                L.kalman_smooth(kal)
                _check(L, kal, _sub(pr, 3, pr.k), 3, oracle, f"{tag}: after kalman_forget(2)")
                last = pr.k - 1
                L.kalman_rollback(kal, last)
                s = eqs.structs[last]
                L.kalman_observe(kal, s.G, s.o, s.C, s.C_type or b"W")
                L.kalman_smooth(kal)
                _check(L, kal, _sub(pr, 3, pr.k), 3, oracle, f"{tag}: after kalman_rollback({last}) and observe")
        finally:
            L.kalman_free(kal)
