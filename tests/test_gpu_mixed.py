"""The mixed-precision odd-even path (FP32 elimination, FP64 iterative
refinement of the estimates, ultimate-kalman_amd/csrc/refine.hip) against
the FP64 odd-even path and the oracle.

Bar: after the refinement steps, per-step normwise relative error of the
estimates <= 1e-9 against the FP64 results (which match the oracle to
~1e-13); each refinement step must shrink the error until it reaches the
FP64 rounding level.
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


def _errors(dev, probs, iters):
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        ref = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
        errs = []
        for it in iters:
            b.smooth_mixed(it)
            e = max(max_rel_err(list(b.download(i, want_cov=False)[0]), list(ref[i])) for i in range(len(probs)))
            errs.append(e)
            say(f"  {it} refinement steps: max error {e:.2e}")
        return errs
    finally:
        b.free()


@pytest.mark.parametrize("case", ["rot6", "n48m24", "n64m32"])
def test_refinement_reaches_fp64(dev, case):
    if case == "rot6":
        F = P.random_orthogonal(P.rng_for(21), 6)
        pr = P.iid_problem(4097, 6, 6, seed=22, F=F, K_kind="W", C_kind="W", name="rot6_k4097")
        iters, final = (0, 1, 2), 2
    elif case == "n48m24":
        F = 0.99 * P.random_orthogonal(P.rng_for(23), 48)
        pr = P.iid_problem(1025, 48, 24, seed=24, F=F, G_scale=1 / np.sqrt(48.0), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name="n48m24_k1025")
        iters, final = (0, 1, 2, 3), 3
    else:  # ill-conditioned: F = Q diag(logspace(0, -3, 64)), K = 1e-2 I
        F = P.random_orthogonal(P.rng_for(25), 64) @ np.diag(np.logspace(0, -3, 64))
        pr = P.iid_problem(257, 64, 32, seed=26, F=F, G_scale=1 / 8.0, Kvar=1e-2, name="n64m32_k257")
        iters, final = (0, 2, 4, 6, 8), 8
    say(f"mixed precision {pr.name}")
    errs = _errors(dev, [pr], iters)
    assert errs[0] > 1e-9  # FP32 alone is not FP64-accurate
    for a, b in zip(errs, errs[1:]):
        assert b < a or b < 1e-13, errs
    assert errs[iters.index(final)] < TOL, errs


def test_mixed_batch_of_dimensions(dev, oracle):
    specs = P.batch_specs(count=24, seed=5, lmin=8, lmax=300)
    probs = [P.batch_problem(s) for s in specs]
    say(f"mixed-precision batch: {len(probs)} problems")
    errs = _errors(dev, probs, (3,))
    assert errs[0] < TOL
    b = dev.Batch(probs)
    b.smooth_mixed(3)
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    try:
        for i, pr in enumerate(probs[:6]):
            x, _ = b.download(i, want_cov=False)
            ox, _ = oracle.oddeven(pr)
            assert max_rel_err(list(x), ox) < TOL, pr.name
    finally:
        oracle.lib.oracle_set_threads(1)
        b.free()


def test_sweep_segments(dev):
    """The refinement sweeps give each step a segment of 16 lanes at n <= 16
    (four steps per wave) and the whole wave above: every dimension class,
    lengths whose levels have odd numbers of even steps and pairs (a wave's
    spare segments), one-step problems. Bars: one refinement step within 1e-9 of the
    FP64 estimates, two within 1e-11. A one-step problem is a single square
    random G (its condition number reaches 1e3 at n = 29, and each step
    multiplies the error by about cond^2 * 6e-8), so it is held to the second
    bar only."""
    probs = []
    for n in (1, 3, 5, 8, 9, 12, 16, 17, 21, 24, 25, 29, 32):
        F = 0.97 * P.random_orthogonal(P.rng_for(700 + n), n)
        for k in (1, 2, 3, 7, 45, 131):
            m = n if k <= 3 else max(1, n // 2)  # every state determined
            probs.append(P.iid_problem(k, n, m, seed=710 + 10 * n + k, F=F, G_scale=1 / np.sqrt(n),
                                       Cvar=0.1, unobserved=lambda i: i % 10 == 9, name=f"n{n}_k{k}"))
    say(f"sweep segments: {len(probs)} problems")
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        ref = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
        errs = {}
        for it in (1, 2):
            b.smooth_mixed(it)
            errs[it] = [max_rel_err(list(b.download(i, want_cov=False)[0]), list(ref[i])) for i in range(len(probs))]
        e1 = max(e for e, pr in zip(errs[1], probs) if pr.k > 1)
        e1_one = max(e for e, pr in zip(errs[1], probs) if pr.k == 1)
        say(f"  one refinement: {e1:.2e} (one-step problems {e1_one:.2e}); two: {max(errs[2]):.2e}")
        assert e1 < TOL and max(errs[2]) < 1e-11, (e1, max(errs[2]))
    finally:
        b.free()


def test_few_observations_stay_finite(dev):
    """m << n (m = 1, 2, 4 with n = 5, 8, 14): the stacked blocks are nearly
    dependent, and the FP32 Householder reflectors must stay within FP32's
    range (a reflector form that squares the column's scale once produced
    NaN here). Two refinement steps reach FP64 accuracy."""
    specs = [dict(n=14, m=4, k=220, seed=101, unobserved=0.05, name="n14m4_k220"),
             dict(n=8, m=2, k=178, seed=102, unobserved=0.05, name="n8m2_k178"),
             dict(n=5, m=1, k=21, seed=103, unobserved=0.05, name="n5m1_k21"),
             dict(n=14, m=4, k=559, seed=104, unobserved=0.05, name="n14m4_k559"),
             dict(n=16, m=1, k=400, seed=105, unobserved=0.05, name="n16m1_k400")]
    probs = [P.serve_problem(s) for s in specs]
    say("few observations per step")
    errs = _errors(dev, probs, (0, 2))
    assert np.isfinite(errs[0]) and errs[1] < TOL, errs


def test_precision_switch_and_covariances(dev, oracle):
    """The precision workload's dynamics (n = 64, m = 32, F = Q
    diag(logspace(0,-3,64)), K = 1e-2 I) at k = 515 against the oracle:
    levels above fp32_levels eliminated in FP64, estimates after refinement
    and the covariance diagonals formed from the same factors."""
    F = P.random_orthogonal(P.rng_for(1004), 64) @ np.diag(np.logspace(0, -3, 64))
    pr = P.iid_problem(515, 64, 32, seed=14, F=F, G_scale=1 / 8.0, Kvar=1e-2, Cvar=1.0, name="n64_illcond_k515")
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    try:
        ox, oc = oracle.oddeven(pr)
    finally:
        oracle.lib.oracle_set_threads(1)
    od = [np.diag(c) for c in oc]
    b = dev.Batch([pr])
    try:
        for fl in (0, 4, 3, 2):
            for r in (1, 2):
                b.smooth_mixed_opt(r, fl, covariances=True)
                x, _ = b.download(0, want_cov=False)
                d = b.download_cov(0, "D")
                ex, ed = max_rel_err(list(x), ox), max_rel_err(list(d), od)
                say(f"  fp32_levels={fl} refinements={r}: estimates {ex:.2e}, covariance diagonals {ed:.2e}")
                if r == 2:
                    assert ex < TOL
                assert ed < 1e-2
    finally:
        b.free()


def test_mixed_graph_replays_match_eager(dev):
    """Mixed-precision smoothing of a batch of several state dimensions,
    captured as one HIP graph (every part on its own forked stream) and
    replayed: each replay, with eager work and synchronizations between them,
    equals the eager launch sequence bit for bit."""
    specs = P.batch_specs(count=12, seed=9, lmin=8, lmax=200)
    probs = [P.batch_problem(s) for s in specs]
    b = dev.Batch(probs)
    L = dev.lib()
    try:
        L.ukhip_use_graphs(0)
        b.smooth_mixed_opt(1, 3, covariances=True)
        eager = [b.download(i) for i in range(len(probs))]
        L.ukhip_use_graphs(1)
        for rep in range(3):
            b.smooth("oddeven")  # other eager work between the replays
            b.smooth_mixed_opt(1, 3, covariances=True)
            for i in range(len(probs)):
                x, c = b.download(i)
                assert np.array_equal(x, eager[i][0]) and np.array_equal(c, eager[i][1]), (rep, i)
    finally:
        L.ukhip_use_graphs(1)
        b.free()


def test_storage_on_demand(dev):
    """A batch smoothed only by the mixed path for its estimates never holds
    the FP64 elimination's storage: asking it for covariances is an error
    (not stale memory). An FP64 smoothing afterwards allocates that storage
    and gives the same result as on a fresh batch; mixed smoothings with other
    precision switches (their FP64 levels' storage made per switch) and with
    covariances then still agree with a fresh batch bit for bit."""
    specs = P.batch_specs(count=6, seed=12, lmin=40, lmax=700)
    probs = [P.batch_problem(s) for s in specs]
    L = dev.lib()
    b = dev.Batch(probs)
    fresh = dev.Batch(probs)
    try:
        b.smooth_mixed_opt(1, 2, False)
        with pytest.raises(dev.UkhipError):
            b.download(0)
        b.smooth("oddeven")
        fresh.smooth("oddeven")
        for i in range(len(probs)):
            x, c = b.download(i)
            xf, cf = fresh.download(i)
            assert np.array_equal(x, xf) and np.array_equal(c, cf), i
        for fl, cov in ((3, False), (1, True), (8, True)):
            b.smooth_mixed_opt(2, fl, cov)
            fresh.smooth_mixed_opt(2, fl, cov)
            for i in range(len(probs)):
                x, c = b.download(i, want_cov=cov)
                xf, cf = fresh.download(i, want_cov=cov)
                assert np.array_equal(x, xf), (fl, cov, i)
                if cov:
                    assert np.array_equal(c, cf), (fl, cov, i)
        say("mixed-only batch: no covariance storage; FP64 and mixed results equal a fresh batch's")
    finally:
        L.ukhip_use_graphs(1)
        b.free()
        fresh.free()
