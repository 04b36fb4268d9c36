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


def _full_inputs(probs, i, n, m, kt, ct):
    """Step i of every problem (random_problem: per-step F, H, c, dense K and C
    in representations kt / ct) as ukhip_filter_step arrays."""
    nf = len(probs)
    F = np.stack([pr.F[i] if pr.F[i] is not None else np.eye(n) for pr in probs])
    H = np.stack([pr.H[i] if pr.H[i] is not None else np.eye(n) for pr in probs])
    c = np.stack([(pr.c[i] if pr.c[i] is not None else np.zeros((n, 1))).ravel() for pr in probs])
    Kd = np.ones((nf, n)) if kt == "w" else np.stack([np.eye(n)] * nf)
    G, o, mi = np.zeros((nf, m, n)), np.zeros((nf, m)), np.zeros(nf, dtype=np.int32)
    Cd = np.ones((nf, m)) if ct == "w" else np.stack([np.eye(m)] * nf)
    for p, pr in enumerate(probs):
        if pr.K[i] is not None:
            Kd[p] = pr.K[i].ravel() if kt == "w" else pr.K[i]
        if pr.o[i] is not None:
            G[p], o[p], mi[p] = pr.G[i], pr.o[i].ravel(), m
            Cd[p] = pr.C[i].ravel() if ct == "w" else pr.C[i]
    return dict(F=F, H=H, c=c, K=Kd, G=G, o=o, C=Cd, m_i=mi, K_type=kt, C_type=ct)


@pytest.mark.parametrize("kt,ct", [("C", "W"), ("W", "C"), ("U", "w"), ("w", "F")])
def test_full_covariance_representations(dev, kt, ct):
    """K and C given as matrices (whitened on the device by k_weigh before the
    filter step), with random H, F, c per step: estimates and information
    matrices against the reference's ULTIMATE filter step by step."""
    nf, n, m, k = 4, 6, 3, 30
    probs = [P.random_problem(950 + p, k, n, m, K_kind=kt, C_kind=ct, H_kind="R",
                              unobserved=tuple(range(p % 3 + 2, k, 7))) for p in range(nf)]
    eqs = [capi.EquationSet(pr) for pr in probs]
    L, kals = _reference_filters(probs)
    flt = dev.Filter(nf, n, m)
    worst_x, worst_info, determined = 0.0, 0.0, 0
    try:
        for i in range(k):
            flt.step(**_full_inputs(probs, i, n, m, kt, ct))
            x, R = flt.estimates()
            for p, (kal, es) in enumerate(zip(kals, eqs)):
                s = es.structs[i]
                L.kalman_evolve(kal, n, s.H, s.F, s.c, s.K, kt.encode())
                L.kalman_observe(kal, s.G, s.o, s.C, ct.encode())
                em, cm = L.kalman_estimate(kal, i), L.kalman_covariance(kal, i)
                rx, rc = capi._read(em)[:, 0], capi._read(cm)
                L.matrix_free(em)
                L.matrix_free(cm)
                if np.all(np.isnan(rx)):
                    assert np.all(np.isnan(x[p])), (p, i)
                    continue
                determined += 1
                worst_x = max(worst_x, rel_err(x[p], rx))
                worst_info = max(worst_info, rel_err(R[p].T @ R[p], rc.T @ rc))
        say(f"K '{kt}', C '{ct}': {determined} determined steps, estimates {worst_x:.2e}, "
            f"information matrices {worst_info:.2e}")
        assert determined > nf * (k - 6) and worst_x < 1e-9 and worst_info < 1e-9
    finally:
        flt.close()
        for kal in kals:
            L.kalman_free(kal)


@pytest.mark.parametrize("kind", ["w", "full"])
def test_smoother_pass_matches_reference(dev, kind):
    """ukhip_filter_smooth over the kept history against the reference's
    ULTIMATE kalman_smooth (c/kalman_ultimate.c:559-620) on every step:
    smoothed estimates, and covariances as information matrices W^T W; and
    against the odd-even smoother on the GPU (covariance = (W^T W)^{-1})."""
    nf, n, m, k = 4, 6, 3, 40
    if kind == "w":
        probs = _problems(nf, n, m, k, unob=True)
        inputs = [_inputs(probs, i, n, m) for i in range(k)]
        kt = ct = "w"
    else:
        kt, ct = "C", "W"
        probs = [P.random_problem(960 + p, k, n, m, K_kind=kt, C_kind=ct, H_kind="R",
                                  unobserved=tuple(range(p % 3 + 2, k, 7))) for p in range(nf)]
        inputs = [_full_inputs(probs, i, n, m, kt, ct) for i in range(k)]
    eqs = [capi.EquationSet(pr) for pr in probs]
    L, kals = _reference_filters(probs)
    L.kalman_smooth.argtypes = [C.c_void_p]
    flt = dev.Filter(nf, n, m)
    try:
        flt.keep_history(k)
        for i in range(k):
            flt.step(**inputs[i])
            for kal, es in zip(kals, eqs):
                s = es.structs[i]
                L.kalman_evolve(kal, n, s.H, s.F, s.c, s.K, kt.encode())
                L.kalman_observe(kal, s.G, s.o, s.C, ct.encode())
        xs, Ws = flt.smooth()
        worst_x, worst_info = 0.0, 0.0
        for p, kal in enumerate(kals):
            L.kalman_smooth(kal)
            for i in range(k):
                em, cm = L.kalman_estimate(kal, i), L.kalman_covariance(kal, i)
                rx, rc = capi._read(em)[:, 0], capi._read(cm)
                assert L.kalman_covariance_type(kal, i) == b"W"
                L.matrix_free(em)
                L.matrix_free(cm)
                worst_x = max(worst_x, rel_err(xs[i, p], rx))
                worst_info = max(worst_info, rel_err(Ws[i, p].T @ Ws[i, p], rc.T @ rc))
        b = dev.Batch(probs if kind == "w" else [P.packed_from_lists(pr) for pr in probs])
        try:
            b.smooth("oddeven")
            worst_oe = 0.0
            for p in range(nf):
                xo, So = b.download(p)
                for i in range(k):
                    worst_oe = max(worst_oe, rel_err(xs[i, p], xo[i]),
                                   rel_err(np.linalg.inv(Ws[i, p].T @ Ws[i, p]), So[i]))
        finally:
            b.free()
        say(f"filter smoother ({kind}): vs reference estimates {worst_x:.2e}, information matrices "
            f"{worst_info:.2e}; vs odd-even {worst_oe:.2e}")
        assert worst_x < 1e-9 and worst_info < 1e-9 and worst_oe < 1e-9
    finally:
        flt.close()
        for kal in kals:
            L.kalman_free(kal)


def test_device_resident_inputs(dev):
    """ukhip_filter_step_device reads the step's arrays in device memory in
    place: the same steps as ukhip_filter_step (host arrays, uploaded) give
    bit-identical estimates, covariance factors and smoother output, including
    unobserved filters (m_i = 0) and a step whose K and C come as matrices
    ('C' and 'W', whitened on the device)."""
    nf, n, m, k = 6, 5, 3, 24
    probs = _problems(nf, n, m, k, True)
    host, devf = dev.Filter(nf, n, m), dev.Filter(nf, n, m)
    host.keep_history(k)
    devf.keep_history(k)
    try:
        for i in range(k):
            kw = _inputs(probs, i, n, m)
            if i == 11:  # matrix representations: K as covariance 'C', C as 'W'
                kw["K"] = np.stack([np.diag(1.0 / np.square(kv)) for kv in kw["K"]])
                kw["C"] = np.stack([np.diag(cv) for cv in kw["C"]])
                kw["K_type"], kw["C_type"] = "C", "W"
            if i == 0:
                kw.pop("F")
            host.step(**kw)
            up = dev.Filter.upload(**kw)
            devf.step_device(up)
            dev.lib().ukhip_synchronize()
            xh, Rh = host.estimates()
            xd, Rd = devf.estimates()
            assert np.array_equal(xh, xd, equal_nan=True) and np.array_equal(Rh, Rd, equal_nan=True), i
            for a in up.values():
                if isinstance(a, dev.DeviceArray):
                    a.free()
        sh, Wh = host.smooth()
        sd, Wd = devf.smooth()
        assert np.array_equal(sh, sd) and np.array_equal(Wh, Wd)
        say(f"device-resident filter inputs: {k} steps bit-identical to host inputs, smoother too")
    finally:
        host.close()
        devf.close()
