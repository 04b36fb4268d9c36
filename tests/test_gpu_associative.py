"""Parity of libukhip's associative-scan smoother (HIP, FP64) with the oracle
and the reference's golden outputs, through the C ABI
(kalman_smooth_associative over kalman_step_equations_t) and the
device-resident batch API.

Bars: against the oracle's associative smoother (same formulas, other
rounding order) per-step normwise relative error <= 1e-9 on estimates and
covariances, or, on problems where the associative algorithm itself departs
from least squares, checks.assoc_parity (no further from the QR-based
odd-even result than the oracle's own associative smoother);
against the reference's stored outputs the oracle's own bar,
1e-7 (the reference's covariance-form combines depart from the global
least-squares solution by up to ~1e-8 on some random problems). The last
step's covariance is the filtered one, which the reference does not return.
"""
import json
import os
import sys

import numpy as np
import pytest

from checks import assoc_parity, max_rel_err, rel_err
from ukhip import capi
from ukhip import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-9
TOL_REF = 1e-7
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    device.lib()
    return device


def test_small_fixture_cases(dev, oracle):
    d = np.load(os.path.join(GOLD, "small_associative.npz"))
    meta = json.loads(bytes(d["meta"]).decode())
    ran = 0
    for tag in meta["cases"]:
        pre = tag + "__"
        a = {k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}
        pr = P.problem_from_arrays(a, tag)
        say(f"small case {tag}")
        xs, cs = dev.smooth_equations(pr, "associative")
        assert max_rel_err(xs, list(a["out_state"])) < TOL_REF, tag
        assert max_rel_err(cs, list(a["out_cov"]), skip=(pr.k - 1,)) < TOL_REF, tag
        ox, oc = oracle.associative(pr)
        ex, ec = max_rel_err(xs, ox), max_rel_err(cs, oc)
        say(f"  vs oracle: state {ex:.2e} cov {ec:.2e}")
        assert ex < TOL and ec < TOL, tag
        ran += 1
    assert ran >= 20


@pytest.mark.parametrize("n,k", [(6, 1024), (48, 128)])
def test_performance_driver_inputs(dev, n, k):
    d = np.load(os.path.join(GOLD, f"performance_n{n}_k{k}_associative.npz"))
    F, G, o, I = d["perf_F"].reshape(n, n), d["perf_G"].reshape(n, n), d["perf_o"].reshape(n, 1), np.eye(n)
    pr = P.Problem(k=k, n=[n] * k, H=[None] + [I] * (k - 1), F=[None] + [F] * (k - 1),
                   c=[None] + [np.zeros((n, 1))] * (k - 1), K=[None] + [I] * (k - 1),
                   K_type=[None] + ["W"] * (k - 1), G=[G] * k, o=[o] * k, C=[I] * k, C_type=["W"] * k)
    say(f"performance inputs n={n} k={k}")
    xs, cs = dev.smooth_equations(pr, "associative")
    ex = max_rel_err(xs, list(d["out_state"]))
    diag = np.stack([np.diag(c) for c in cs])
    ed = rel_err(diag[:k - 1], d["out_covdiag"][:k - 1])
    say(f"  vs reference: state {ex:.2e} cov diag {ed:.2e}")
    assert ex < TOL_REF and ed < TOL_REF
    for j, i in enumerate(d["out_cov_sel_index"]):
        if i != k - 1:
            assert rel_err(cs[i], d["out_cov_sel"][j]) < TOL_REF


def _compare_batch(dev, oracle, probs, tol=TOL):
    b = dev.Batch(probs)
    b.smooth("associative")
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    try:
        for i, pr in enumerate(probs):
            x, cv = b.download(i)
            ox, oc = oracle.associative(pr)
            ok, what = assoc_parity((list(x), list(cv)), (ox, oc), oracle.oddeven(pr), tol=tol)
            say(f"  problem {i} ({pr.name}, k={pr.k}): {what}")
            assert ok, pr.name
    finally:
        oracle.lib.oracle_set_threads(1)
        b.free()


@pytest.mark.parametrize("case", ["rot6", "n48m24", "n64m32"])
def test_configuration_shapes(dev, oracle, case):
    if case == "rot6":
        F = P.random_orthogonal(P.rng_for(21), 6)
        pr = P.iid_problem(4097, 6, 6, seed=22, F=F, K_kind="W", C_kind="W", name="rot6_k4097")
    elif case == "n48m24":  # m < n: the prior spans steps 0 and 1
        F = 0.99 * P.random_orthogonal(P.rng_for(23), 48)
        pr = P.iid_problem(1025, 48, 24, seed=24, F=F, G_scale=1 / np.sqrt(48.0), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name="n48m24_k1025")
    else:
        F = P.random_orthogonal(P.rng_for(25), 64) @ np.diag(np.logspace(0, -3, 64))
        pr = P.iid_problem(257, 64, 32, seed=26, F=F, G_scale=1 / 8.0, Kvar=1e-2, name="n64m32_k257")
    say(f"configuration shape {pr.name}")
    _compare_batch(dev, oracle, [pr])


def test_batch_of_mixed_lengths(dev, oracle):
    """Uneven lengths in one batch, unobserved steps, m < n (prior over
    several steps) and k = 1."""
    F = P.random_orthogonal(P.rng_for(31), 8)
    probs = [P.uniform_problem(k, 8, 4, seed=40 + k, F=0.97 * F,
                               G_fn=lambda r, i: r.standard_normal((4, 8)) / np.sqrt(8.0),
                               unobserved=(lambda i: i % 7 == 3), name=f"mix_k{k}")
             for k in (2, 3, 5, 64, 1000, 777)]
    probs.insert(0, P.uniform_problem(1, 8, 8, seed=39, F=F, name="mix_k1"))
    say("batch of mixed lengths")
    _compare_batch(dev, oracle, probs)


def test_general_H_and_covariance_types(dev, oracle):
    """H != I, covariance types C/U/F/W/w, unobserved steps inside the prior
    window (the first steps do not determine the state)."""
    probs = [P.random_problem(81, 40, 4, 2, K_kind="C", C_kind="w", H_kind="R", unobserved=(1, 2, 17)),
             P.random_problem(82, 33, 5, 3, K_kind="U", C_kind="F", unobserved=(0, 4)),
             P.random_problem(83, 50, 3, 1, K_kind="W", C_kind="C", H_kind="R")]
    for pr in probs:
        say(f"general case {pr.name}")
        xs, cs = dev.smooth_equations(pr, "associative")
        ox, oc = oracle.associative(pr)
        ex, ec = max_rel_err(xs, ox), max_rel_err(cs, oc)
        say(f"  state err {ex:.2e} cov err {ec:.2e}")
        assert ex < TOL and ec < TOL, pr.name


def test_rotation_fixture_k100000(dev):
    """The reference's associative outputs at k = 100,000 (n = 6 rotation),
    every 1024th step, plus whole-run summaries; the interior covariance is
    (1/sqrt 5) I."""
    d = np.load(os.path.join(GOLD, "rot6_k100000_associative.npz"))
    meta = json.loads(bytes(d["meta"]).decode())
    g = dict(meta["gen"])
    pr = P.iid_problem(meta["k"], meta["n"], meta["m"], F=d["gen_F"], **g)
    if P.input_digest(pr) != meta["sha256"]:
        pytest.skip("regenerated inputs differ from the fixture's")
    say("rotation k=100000 associative")
    b = dev.Batch([pr])
    b.smooth("associative")
    x, cv = b.download(0)
    b.free()
    idx = d["out_index"]
    assert max_rel_err([x[i] for i in idx], list(d["out_state"])) < TOL_REF
    assert rel_err(np.max(np.abs(x), axis=0), d["out_state_maxabs"]) < TOL_REF
    assert rel_err(np.sum(x ** 2, axis=0), d["out_state_sumsq"]) < TOL_REF
    diag = np.stack([np.diag(cv[i]) for i in idx])
    keep = idx != pr.k - 1
    assert rel_err(diag[keep], d["out_covdiag"][keep]) < TOL_REF
    assert np.max(np.abs(cv[40:-40] - np.eye(6) / np.sqrt(5))) < 1e-12


def test_agrees_with_oddeven(dev):
    """Both smoothers compute the same least-squares estimates and
    covariances (the last covariance: filtered = smoothed at the last step)."""
    F = 0.99 * P.random_orthogonal(P.rng_for(91), 16)
    pr = P.iid_problem(3001, 16, 8, seed=92, F=F, G_scale=0.25, Cvar=0.1,
                       unobserved=lambda i: i % 10 == 9, name="n16m8_k3001")
    b = dev.Batch([pr])
    b.smooth("oddeven")
    xo, co = b.download(0)
    b.smooth("associative")
    xa, ca = b.download(0)
    b.free()
    ex, ec = max_rel_err(list(xa), list(xo)), max_rel_err(list(ca), list(co))
    say(f"associative vs odd-even: state {ex:.2e} cov {ec:.2e}")
    assert ex < 1e-9 and ec < 1e-9


def test_kalman_t_associative(dev, oracle):
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
    pr = P.random_problem(72, 29, 3, 3, K_kind="W", C_kind="C", unobserved=(5, 6))
    eqs = capi.EquationSet(pr)
    kal = L.kalman_create_options(capi.KALMAN_ALGORITHM_ASSOCIATIVE)
    assert kal
    for i in range(pr.k):
        s = eqs.structs[i]
        L.kalman_evolve(kal, pr.n[i], s.H, s.F, s.c, s.K, s.K_type or b"W")
        L.kalman_observe(kal, s.G, s.o, s.C, s.C_type or b"W")
    say("kalman_t associative smooth")
    L.kalman_smooth(kal)
    ox, oc = oracle.associative(pr)
    for i in range(pr.k):
        e = L.kalman_estimate(kal, i)
        cvm = L.kalman_covariance(kal, i)
        x = capi._read(e)[:, 0]
        cv = capi._read(cvm)
        L.matrix_free(e)
        L.matrix_free(cvm)
        assert rel_err(x, ox[i]) < TOL and rel_err(cv, oc[i]) < TOL, i
    L.kalman_free(kal)


def test_timed_classes(dev):
    """ukhip_batch_smooth_associative_timed: the same results as the untimed
    call, bit for bit, on a batch of two state dimensions; every kernel class
    launched, the class times adding up to the span."""
    probs = [P.iid_problem(k, n, n // 2 + 1, seed=700 + k, F=0.97 * P.random_orthogonal(P.rng_for(701 + n), n),
                           G_scale=1 / np.sqrt(n), name=f"n{n}_k{k}") for n, k in ((6, 300), (20, 257))]
    b = dev.Batch(probs)
    try:
        b.smooth("associative")
        ref = [b.download(i) for i in range(len(probs))]
        ms, nl = b.smooth_associative_timed()
        for i in range(len(probs)):
            x, S = b.download(i)
            assert np.array_equal(x, ref[i][0]) and np.array_equal(S, ref[i][1]), i
        classes = [c for c in dev.AS_KERNELS if c != "total"]
        assert all(nl[c] >= 2 for c in classes), nl  # two parts, every class launched in each
        assert abs(sum(ms[c] for c in classes) - ms["total"]) < 1e-3 * ms["total"] + 1e-3, ms
        say("associative classes: " + ", ".join(f"{c} {ms[c]:.3f} ms" for c in dev.AS_KERNELS))
    finally:
        b.free()
