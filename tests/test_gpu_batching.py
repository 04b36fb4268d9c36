"""Batches of many independent problems of several state dimensions (the
batching workload's shape at a size the oracle checks in seconds), smoothed
by replaying a captured HIP graph.

Bars: odd-even, each problem within 1e-9 (per-step normwise relative error)
of the oracle; associative, checks.assoc_parity (1e-9 of the oracle, or as
close to the least-squares result as the oracle's own associative smoother);
graph replays bit-identical to the eager launch sequence and to each other,
with eager work and synchronisations between the replays.
"""
import os
import sys

import numpy as np
import pytest

from checks import assoc_parity, max_rel_err
from ukhip import capi
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


@pytest.fixture(scope="module")
def mixed():
    specs = P.batch_specs(count=48, seed=7, lmin=8, lmax=400)
    return [P.batch_problem(s) for s in specs]


def _check_all(b, probs, oracle, algorithm, options=0):
    want_cov = algorithm == "associative" or not (options & capi.KALMAN_NO_COVARIANCE)
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    worst = 0.0
    try:
        for i, pr in enumerate(probs):
            x, cv = b.download(i, want_cov)
            ox, oc = getattr(oracle, algorithm)(pr, options)
            if algorithm == "associative":
                ok, what = assoc_parity((list(x), list(cv)), (ox, oc), oracle.oddeven(pr))
                if "least squares" in what:
                    say(f"  {pr.name}: {what}")
                assert ok, (pr.name, what)
                continue
            e = max_rel_err(list(x), ox)
            if want_cov:
                e = max(e, max_rel_err(list(cv), oc))
            worst = max(worst, e)
            assert e < TOL, (pr.name, e)
    finally:
        oracle.lib.oracle_set_threads(1)
    return worst


def test_mixed_dimensions_graph(dev, oracle, mixed):
    dims = sorted({p.n[0] for p in mixed})
    say(f"mixed batch: {len(mixed)} problems, n in {dims}, {sum(p.k for p in mixed)} steps")
    b = dev.Batch(mixed)
    try:
        dev.lib().ukhip_use_graphs(1)
        b.smooth("oddeven")
        worst = _check_all(b, mixed, oracle, "oddeven")
        say(f"  graph replay vs oracle: worst {worst:.2e}")
    finally:
        b.free()


def test_graph_replays_match_eager(dev, mixed):
    L = dev.lib()
    b = dev.Batch(mixed)
    try:
        L.ukhip_use_graphs(0)
        b.smooth("oddeven")
        ref = [b.download(i) for i in range(len(mixed))]
        L.ukhip_use_graphs(1)
        for rep in range(3):
            b.smooth("oddeven")
            got = [b.download(i) for i in range(len(mixed))]
            for (x0, c0), (x1, c1) in zip(ref, got):
                assert np.array_equal(x0, x1) and np.array_equal(c0, c1), rep
            L.ukhip_use_graphs(0)  # eager work between replays
            b.smooth("oddeven", options=capi.KALMAN_NO_COVARIANCE)
            L.ukhip_use_graphs(1)
            say(f"  replay {rep} identical to eager")
        # the no-covariance graph is captured separately; its estimates are the
        # covariance path's to rounding (for n > 16 another kernel computes them)
        b.smooth("oddeven", options=capi.KALMAN_NO_COVARIANCE)
        for i in range(len(mixed)):
            x, _ = b.download(i, want_cov=False)
            assert max_rel_err(list(x), list(ref[i][0])) < 1e-12
    finally:
        L.ukhip_use_graphs(1)
        b.free()


def test_mixed_dimensions_associative(dev, oracle, mixed):
    sub = mixed[:24]
    b = dev.Batch(sub)
    try:
        b.smooth("associative")
        worst = _check_all(b, sub, oracle, "associative")
        say(f"  associative mixed batch vs oracle: worst {worst:.2e}")
    finally:
        b.free()


def test_shared_F_is_the_expanded_F(dev):
    """F given once (F_shared) and F given per step give identical results."""
    spec = P.batch_specs(count=1, seed=11, lmin=300, lmax=301)[0]
    pr = P.batch_problem(spec)
    expanded = P.batch_problem(spec)
    expanded.packed = dict(expanded.packed)
    expanded.packed["F_shared"] = None
    out = []
    for q in (pr, expanded):
        b = dev.Batch([q])
        b.smooth("oddeven")
        out.append(b.download(0))
        b.free()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_many_problems_one_dimension(dev, oracle):
    """4,200 problems of one state dimension in one batch: at level 0 the
    kernels' 64-way problem search runs three rounds (beyond 64 * 64
    problems), and every unit must land in its own problem. Lengths cycle
    through 1..9 with a longer problem every 97th; a sample spread over the
    whole index range is checked against the oracle, every result for
    finiteness."""
    n = 5
    F = 0.97 * P.random_orthogonal(P.rng_for(1201), n)
    probs = []
    for i in range(4200):
        k = 40 if i % 97 == 0 else 1 + i % 9
        probs.append(P.iid_problem(k, n, n, seed=1300 + i, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                                   name=f"p{i}_k{k}"))
    say(f"one dimension, {len(probs)} problems, {sum(p.k for p in probs)} steps")
    b = dev.Batch(probs)
    try:
        b.smooth("oddeven")
        for i in range(len(probs)):
            x, _ = b.download(i, want_cov=False)
            assert np.all(np.isfinite(x)), probs[i].name
        sample = [probs[i] for i in range(0, len(probs), 71)] + [probs[-1]]
        idx = list(range(0, len(probs), 71)) + [len(probs) - 1]
        sub_b = _Sub(b, idx)
        worst = _check_all(sub_b, sample, oracle, "oddeven")
        say(f"  {len(sample)} sampled problems vs oracle: worst {worst:.2e}")
    finally:
        b.free()


class _Sub:
    """A view of some problems of a batch, for _check_all."""

    def __init__(self, b, idx):
        self.b, self.idx = b, idx

    def download(self, i, want_cov=True):
        return self.b.download(self.idx[i], want_cov)
