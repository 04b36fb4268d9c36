"""Parity of libukhip's odd-even smoother (HIP, FP64) with the oracle and the
reference's golden outputs, through the C ABI (kalman_smooth_oddeven over
kalman_step_equations_t) and the device-resident batch API.

Bar (FP64 path): per-step normwise relative error <= 1e-9 on estimates and
covariances; observed differences are ~1e-13.
"""
import json
import os
import sys

import numpy as np
import pytest

import uk_libs
from checks import max_rel_err, rel_err
from ukhip import capi
from ukhip import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    device.lib()
    return device


def _oracle_threads(o):
    o.lib.oracle_set_threads(min(16, os.cpu_count() or 1))


def test_small_fixture_cases(dev, oracle):
    d = np.load(os.path.join(GOLD, "small_oddeven.npz"))
    meta = json.loads(bytes(d["meta"]).decode())
    ran = 0
    for tag in meta["cases"]:
        pre = tag + "__"
        a = {k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}
        pr = P.problem_from_arrays(a, tag)
        say(f"small case {tag}")
        xs, cs = dev.smooth_equations(pr, "oddeven")
        assert max_rel_err(xs, list(a["out_state"])) < TOL, tag
        assert max_rel_err(cs, list(a["out_cov"])) < TOL, tag
        ox, oc = oracle.oddeven(pr)
        assert max_rel_err(xs, ox) < TOL and max_rel_err(cs, oc) < TOL, tag
        ran += 1
    assert ran == len(meta["cases"]) and ran >= 80  # every case, m > n included


def _shape_outputs(a, pr):
    xs, cs, p, q = [], [], 0, 0
    for n in pr.n:
        xs.append(a["out_state_flat"][p:p + n])
        cs.append(a["out_cov_flat"][q:q + n * n].reshape(n, n, order="F"))
        p += n
        q += n * n
    return xs, cs


def test_shape_fixture_cases(dev, oracle):
    """The reference's shape catalogue (kalmanTest.m:80-95 plus larger random
    tables): rectangular H_i, state dimensions that change from step to step,
    more observation rows than states. Against the reference's outputs where
    it ran and against the oracle on every case, with covariances and in the
    'W' representation."""
    d = np.load(os.path.join(GOLD, "shapes_oddeven.npz"))
    meta = json.loads(bytes(d["meta"]).decode())
    for tag in meta["cases"]:
        pre = tag + "__"
        a = {k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}
        pr = P.problem_from_arrays(a, tag)
        say(f"shape case {tag} k={pr.k} n={min(pr.n)}..{max(pr.n)}")
        xs, cs = dev.smooth_equations(pr, "oddeven")
        ox, oc = oracle.oddeven(pr)
        assert max_rel_err(xs, ox) < TOL and max_rel_err(cs, oc) < TOL, tag
        if tag in meta["reference_ok"]:
            rx, rc = _shape_outputs(a, pr)
            assert max_rel_err(xs, rx) < TOL and max_rel_err(cs, rc) < TOL, tag
        x2, _ = dev.smooth_equations(pr, "oddeven", options=capi.KALMAN_NO_COVARIANCE)
        assert max_rel_err(x2, ox) < TOL, tag
        xw, ws = dev.smooth_equations(pr, "oddeven", options=capi.UKHIP_COVARIANCE_W)
        assert max_rel_err(xw, ox) < TOL, tag
        worst = max(rel_err(w @ c @ w.T, np.eye(len(c))) for w, c in zip(ws, oc))
        assert worst < TOL, tag


def test_associative_rejects_rectangular_h(dev):
    """The associative elements need square H and one dimension (as the
    reference's, which never reads H): such a problem fails with an error."""
    d = np.load(os.path.join(GOLD, "shapes_oddeven.npz"))
    tag = "u3238_UW"
    pre = tag + "__"
    pr = P.problem_from_arrays({k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}, tag)
    with pytest.raises(dev.UkhipError, match="associative smoother needs"):
        dev.smooth_equations(pr, "associative")


def test_batch_wide_observations(dev, oracle):
    """The batch API with more observation rows than states (m = n + 3 and
    per-step m_i up to it), FP64 with covariances and the mixed-precision
    path, against the oracle."""
    pr = P.packed_from_lists(P.random_problem(91, 33, 5, 8, K_kind="w", C_kind="w",
                                              m_of=lambda i: (5, 3, 8, 0, 6)[i % 5]))
    pr2 = P.packed_from_lists(P.random_problem(92, 20, 5, 7, K_kind="C", C_kind="W", H_kind="R"))
    say("batch with m > n")
    b = dev.Batch([pr, pr2])
    b.smooth("oddeven")
    for i, q in enumerate((pr, pr2)):
        x, S = b.download(i)
        ox, oc = oracle.oddeven(q)
        assert max_rel_err(list(x), ox) < TOL and max_rel_err(list(S), oc) < TOL
    b.smooth_mixed(2)
    for i, q in enumerate((pr, pr2)):
        x, _ = b.download(i, want_cov=False)
        ox, _ = oracle.oddeven(q)
        assert max_rel_err(list(x), ox) < 1e-8
    b.free()


@pytest.mark.parametrize("n,k", [(6, 1024), (48, 128)])
def test_performance_driver_inputs(dev, n, k):
    d = np.load(os.path.join(GOLD, f"performance_n{n}_k{k}_oddeven.npz"))
    F, G, o, I = d["perf_F"].reshape(n, n), d["perf_G"].reshape(n, n), d["perf_o"].reshape(n, 1), np.eye(n)
    pr = P.Problem(k=k, n=[n] * k, H=[None] + [I] * (k - 1), F=[None] + [F] * (k - 1),
                   c=[None] + [np.zeros((n, 1))] * (k - 1), K=[None] + [I] * (k - 1),
                   K_type=[None] + ["W"] * (k - 1), G=[G] * k, o=[o] * k, C=[I] * k, C_type=["W"] * k)
    say(f"performance inputs n={n} k={k}")
    xs, cs = dev.smooth_equations(pr, "oddeven")
    assert max_rel_err(xs, list(d["out_state"])) < TOL
    assert rel_err(np.stack([np.diag(c) for c in cs]), d["out_covdiag"]) < TOL
    for j, i in enumerate(d["out_cov_sel_index"]):
        assert rel_err(cs[i], d["out_cov_sel"][j]) < TOL


def _compare_batch(dev, oracle, probs, options=0, algorithm="oddeven"):
    b = dev.Batch(probs)
    b.smooth(algorithm, options)
    want_cov = not (options & capi.KALMAN_NO_COVARIANCE)
    _oracle_threads(oracle)
    try:
        for i, pr in enumerate(probs):
            x, cv = b.download(i, want_cov)
            ox, oc = getattr(oracle, algorithm)(pr, options)
            ex = max_rel_err(list(x), ox)
            say(f"  problem {i} ({pr.name}, k={pr.k}): state err {ex:.2e}")
            assert ex < TOL, pr.name
            if want_cov:
                ec = max_rel_err(list(cv), oc)
                say(f"  problem {i}: cov err {ec:.2e}")
                assert ec < TOL, pr.name
    finally:
        oracle.lib.oracle_set_threads(1)
        b.free()


@pytest.mark.parametrize("case", ["rot6", "n48m24", "n64m32"])
def test_configuration_shapes(dev, oracle, case):
    if case == "rot6":
        F = P.random_orthogonal(P.rng_for(21), 6)
        pr = P.iid_problem(4097, 6, 6, seed=22, F=F, K_kind="W", C_kind="W", name="rot6_k4097")
    elif case == "n48m24":
        F = 0.99 * P.random_orthogonal(P.rng_for(23), 48)
        pr = P.iid_problem(1025, 48, 24, seed=24, F=F, G_scale=1 / np.sqrt(48.0), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name="n48m24_k1025")
    else:
        F = P.random_orthogonal(P.rng_for(25), 64) @ np.diag(np.logspace(0, -3, 64))
        pr = P.iid_problem(257, 64, 32, seed=26, F=F, G_scale=1 / 8.0, Kvar=1e-2, name="n64m32_k257")
    say(f"configuration shape {pr.name}")
    _compare_batch(dev, oracle, [pr])


def test_batch_of_mixed_lengths(dev, oracle):
    F = P.random_orthogonal(P.rng_for(31), 8)
    probs = [P.uniform_problem(k, 8, 4, seed=40 + k, F=0.97 * F,
                               G_fn=lambda r, i: r.standard_normal((4, 8)) / np.sqrt(8.0),
                               unobserved=(lambda i: i % 7 == 3), name=f"mix_k{k}")
             for k in (1, 2, 3, 5, 64, 1000, 777)]
    probs[0] = P.uniform_problem(1, 8, 8, seed=39, F=F, name="mix_k1")  # one step needs m = n
    say("batch of mixed lengths")
    _compare_batch(dev, oracle, probs)


def test_no_covariance(dev, oracle):
    F = P.random_orthogonal(P.rng_for(51), 16)
    pr = P.uniform_problem(301, 16, 8, seed=52, F=0.98 * F,
                           G_fn=lambda r, i: r.standard_normal((8, 16)) / 4.0, name="nocov_n16")
    say("no covariance")
    _compare_batch(dev, oracle, [pr], options=capi.KALMAN_NO_COVARIANCE)
    xs, cs = dev.smooth_equations(pr, "oddeven", capi.KALMAN_NO_COVARIANCE)
    assert all(c is None for c in cs)


def test_rotation_closed_form_full_size(dev):
    """k = 100,000, n = 6, F orthogonal, H = G = K = C = I: interior smoothed
    covariance (1/sqrt 5) I and endpoints ((sqrt 5 - 1)/2) I."""
    F = P.random_orthogonal(P.rng_for(61), 6)
    pr = P.iid_problem(100_000, 6, 6, seed=62, F=F, K_kind="w", C_kind="w", name="rot6_k100000")
    say("closed form k=100000")
    b = dev.Batch([pr])
    b.smooth("oddeven")
    x, cv = b.download(0)
    b.free()
    mid = cv[40:-40]  # the boundary layer decays geometrically; 40 steps in it is below 1e-15
    assert np.max(np.abs(mid - np.eye(6) / np.sqrt(5))) < 1e-12
    end = (np.sqrt(5) - 1) / 2
    assert np.max(np.abs(cv[0] - end * np.eye(6))) < 1e-12
    assert np.max(np.abs(cv[-1] - end * np.eye(6))) < 1e-12
    assert np.all(np.isfinite(x))


def test_kalman_t_api(dev, oracle):
    """kalman_create_options / evolve / observe / smooth / estimate / covariance."""
    import ctypes as C
    L = dev.lib()
    L.kalman_create_options.restype = C.c_void_p
    L.kalman_create_options.argtypes = [C.c_uint32]
    L.kalman_evolve.argtypes = [C.c_void_p, C.c_int32] + [C.POINTER(capi.KalmanMatrix)] * 4 + [C.c_char]
    L.kalman_observe.argtypes = [C.c_void_p] + [C.POINTER(capi.KalmanMatrix)] * 3 + [C.c_char]
    L.kalman_smooth.argtypes = [C.c_void_p]
    L.kalman_estimate.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_estimate.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_covariance.argtypes = [C.c_void_p, C.c_int32]
    L.kalman_covariance.restype = C.POINTER(capi.KalmanMatrix)
    L.kalman_free.argtypes = [C.c_void_p]
    pr = P.random_problem(71, 37, 3, 2, K_kind="C", C_kind="w", unobserved=(5, 6))
    eqs = capi.EquationSet(pr)  # borrow its matrix structs
    kal = L.kalman_create_options(capi.KALMAN_ALGORITHM_ODDEVEN)
    assert kal
    for i in range(pr.k):
        s = eqs.structs[i]
        L.kalman_evolve(kal, pr.n[i], s.H, s.F, s.c, s.K, s.K_type or b"W")
        L.kalman_observe(kal, s.G, s.o, s.C, s.C_type or b"W")
    say("kalman_t smooth")
    L.kalman_smooth(kal)
    ox, oc = oracle.oddeven(pr)
    for i in range(pr.k):
        e = L.kalman_estimate(kal, i)
        cvm = L.kalman_covariance(kal, i)
        x = capi._read(e)[:, 0]
        cv = capi._read(cvm)
        L.matrix_free(e)
        L.matrix_free(cvm)
        assert rel_err(x, ox[i]) < TOL and rel_err(cv, oc[i]) < TOL, i
    L.kalman_free(kal)


@pytest.mark.parametrize("H_kind", ["I", "near-I"])
def test_diagonal_weights_with_and_without_identity_H(dev, oracle, H_kind):
    """Every K and C given as 'w' at n = 40 (bucket 48), m = 20: with H = I the
    level-0 D blocks are diagonal and phase 2 skips their zeros; with a dense
    H (the identity plus a small random part, well conditioned) they are dense
    and it must not. Both against the oracle."""
    _oracle_threads(oracle)
    pr = P.random_problem(91, 65, 40, 20, K_kind="w", C_kind="w", H_kind="I", unobserved=(7, 30))
    if H_kind != "I":
        rng = P.rng_for(94)
        for i in range(1, pr.k):
            pr.H[i] = np.eye(40) + 0.05 * rng.standard_normal((40, 40))
    say(f"n=40 m=20 'w' covariances, H={H_kind}")
    xs, cs = dev.smooth_equations(pr, "oddeven")
    ox, oc = oracle.oddeven(pr)
    assert max_rel_err(xs, ox) < TOL and max_rel_err(cs, oc) < TOL


@pytest.mark.parametrize("n", [17, 20, 24])
def test_bucket24(dev, oracle, n):
    """State dimensions 17..24 run the elimination and the refinement sweeps
    in bucket 24 (uk_bucket_qr), whose half-block structure keeps 16 rows:
    m = 16 takes the half-block kernels, m = 17 and m = n the dense ones (one
    batch per m: the structure is chosen per batch part). FP64 estimates and
    covariances against the oracle; the mixed path after two refinement steps
    against the FP64 estimates; 'w' and 'C' weights."""
    F = 0.97 * P.random_orthogonal(P.rng_for(960 + n), n)
    _oracle_threads(oracle)
    worst, worst_mixed = 0.0, 0.0
    try:
        for m, kind in ((5, "w"), (16, "C"), (min(17, n), "w"), (n, "C")):
            probs = [P.iid_problem(k, n, m, seed=970 + k + m, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1, K_kind=kind,
                                   C_kind=kind, unobserved=lambda i: i % 10 == 9, name=f"n{n}_m{m}_{kind}_k{k}")
                     for k in ((3, 130, 517) if 3 * m >= n else (7, 130, 517))]  # k m >= n: a unique solution
            b = dev.Batch(probs)
            try:
                b.smooth("oddeven")
                fp64 = [b.download(i) for i in range(len(probs))]
                b.smooth_mixed_opt(2, 0, False)
                for i, pr in enumerate(probs):
                    ox, oc = oracle.oddeven(pr)
                    e = max(max_rel_err(list(fp64[i][0]), ox), max_rel_err(list(fp64[i][1]), oc))
                    em = max_rel_err(list(b.download(i, want_cov=False)[0]), list(fp64[i][0]))
                    worst, worst_mixed = max(worst, e), max(worst_mixed, em)
                    assert e < TOL and em < TOL, pr.name
            finally:
                b.free()
        say(f"n={n}: FP64 vs oracle {worst:.1e}, mixed (2 refinements) vs FP64 {worst_mixed:.1e}")
    finally:
        oracle.lib.oracle_set_threads(1)
