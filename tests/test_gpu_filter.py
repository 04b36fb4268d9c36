"""The batched sequential UltimateKalman filter (ukhip_filter_*).

Checked step by step against the reference's own KALMAN_ALGORITHM_ULTIMATE
filter (oracle/_ref, the reference C built from its sources:
c/kalman_ultimate.c:140-470 through kalman_evolve / kalman_observe /
kalman_estimate / kalman_covariance), and against the smoothers: at a
problem's last step the filtered estimate and covariance are the smoothed
ones (the odd-even smoother on the GPU).

Bars: estimates within 1e-9 (normwise relative) and NaN exactly where the
reference's are NaN (state not yet determined); covariances compared as
information matrices R^T R (the rows of a QR factor have arbitrary signs),
1e-9.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from checks import max_rel_err, rel_err
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


def _problems(nf, n, m, k, unob):
    probs = []
    for p in range(nf):
        F = 0.98 * P.random_orthogonal(P.rng_for(900 + p), n)
        probs.append(P.iid_problem(k, n, m, seed=910 + p, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                                   unobserved=(lambda i, p=p: (i + p) % 7 == 3) if unob else None,
                                   name=f"filter{p}_n{n}_m{m}"))
    return probs


def _inputs(probs, i, n, m):
    """Step i of every problem as ukhip_filter_step arrays."""
    nf = len(probs)
    F = np.stack([pr.F[i] if pr.F[i] is not None else np.eye(n) for pr in probs])
    K = np.stack([(pr.K[i] if pr.K[i] is not None else np.ones((n, 1))).ravel() for pr in probs])
    G = np.zeros((nf, m, n))
    o = np.zeros((nf, m))
    Cw = np.ones((nf, m))
    mi = np.zeros(nf, dtype=np.int32)
    for p, pr in enumerate(probs):
        if pr.o[i] is not None:
            G[p], o[p], Cw[p], mi[p] = pr.G[i], pr.o[i].ravel(), pr.C[i].ravel(), m
    return dict(F=F, K=K, G=G, o=o, C=Cw, m_i=mi)


def _reference_filters(probs):
    import uk_libs
    ref = uk_libs.reference()
    if ref is None:
        pytest.skip("the reference build (oracle/_ref) is absent")
    L = ref.lib
    L.kalman_create_options.restype = C.c_void_p
    L.kalman_create_options.argtypes = [C.c_uint32]
    L.kalman_evolve.argtypes = [C.c_void_p, C.c_int32] + [C.POINTER(capi.KalmanMatrix)] * 4 + [C.c_char]
    L.kalman_observe.argtypes = [C.c_void_p] + [C.POINTER(capi.KalmanMatrix)] * 3 + [C.c_char]
    L.kalman_estimate.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_estimate.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_covariance.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_covariance.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_covariance_type.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_covariance_type.restype = C.c_char
    L.kalman_free.argtypes = [C.c_void_p]
    L.matrix_free.argtypes = [C.POINTER(capi.KalmanMatrix)]
    kals = [L.kalman_create_options(1) for _ in probs]  # KALMAN_ALGORITHM_ULTIMATE (c/kalman.h:82)
    assert all(kals)
    return L, kals


def test_matches_reference_ultimate_step_by_step(dev):
    nf, n, m, k = 6, 6, 2, 40
    probs = _problems(nf, n, m, k, unob=True)
    eqs = [capi.EquationSet(pr) for pr in probs]
    L, kals = _reference_filters(probs)
    flt = dev.Filter(nf, n, m)
    worst_x, worst_info, determined = 0.0, 0.0, 0
    try:
        for i in range(k):
            flt.step(**_inputs(probs, i, n, m))
            x, R = flt.estimates()
            for p, (kal, es) in enumerate(zip(kals, eqs)):
                s = es.structs[i]
                L.kalman_evolve(kal, n, s.H, s.F, s.c, s.K, b"w")
                L.kalman_observe(kal, s.G, s.o, s.C, b"w")
                em = L.kalman_estimate(kal, i)
                cm = L.kalman_covariance(kal, i)
                rx = capi._read(em)[:, 0]
                rc = capi._read(cm)
                assert L.kalman_covariance_type(kal, i) == b"W"
                L.matrix_free(em)
                L.matrix_free(cm)
                if np.all(np.isnan(rx)):
                    assert np.all(np.isnan(x[p])), (p, i)
                    continue
                determined += 1
                worst_x = max(worst_x, rel_err(x[p], rx))
                worst_info = max(worst_info, rel_err(R[p].T @ R[p], rc.T @ rc))
                assert np.array_equal(R[p], np.triu(R[p]))
        say(f"{nf} filters x {k} steps ({determined} determined): estimates {worst_x:.2e}, "
            f"information matrices {worst_info:.2e}")
        assert determined > nf * (k - 6) and worst_x < 1e-9 and worst_info < 1e-9
    finally:
        flt.close()
        for kal in kals:
            L.kalman_free(kal)


def test_last_step_is_the_smoothed_step(dev):
    """Filtered = smoothed at the last step (the smoother conditions on the
    same observations there)."""
    nf, n, m, k = 5, 16, 8, 300
    probs = _problems(nf, n, m, k, unob=True)
    flt = dev.Filter(nf, n, m)
    try:
        for i in range(k):
            flt.step(**_inputs(probs, i, n, m))
        x, R = flt.estimates()
    finally:
        flt.close()
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        for p in range(nf):
            xs, Ss = b.download(p)
            ex = rel_err(x[p], xs[-1])
            eS = rel_err(np.linalg.inv(R[p].T @ R[p]), Ss[-1])
            say(f"  filter {p}: last estimate {ex:.2e}, last covariance {eS:.2e}")
            assert ex < 1e-9 and eS < 1e-9
    finally:
        b.free()
