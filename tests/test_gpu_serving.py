"""The serving loop (ukhip_server_*): continuous batching of smoothing
requests under a device-memory budget.

Bars: a request's result is bit-identical alone and in a mix of other
requests, and bit-identical to the same problem smoothed through the batch
API (the kernels treat each problem independently); results match the oracle
to 1e-9; the mixed-precision server's estimates are within 1e-9 of FP64; a
request larger than the budget fails with an error instead of stalling.
"""
import os
import random
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


@pytest.fixture(scope="module")
def requests_():
    specs = P.serve_specs(40, seed=17, lmin=20, lmax=600)
    return [P.serve_problem(s) for s in specs]


def test_alone_and_in_the_mix(dev, oracle, requests_):
    probe = requests_[7]
    srv = dev.Server(hbm_budget_bytes=2 << 30, max_batch=8, max_inflight=3)
    try:
        x1, c1, lat1 = srv.wait(srv.submit(probe))
        say(f"alone: {probe.name} latency {lat1:.2f} ms")
        order = list(range(len(requests_)))
        random.Random(3).shuffle(order)
        ids = {i: srv.submit(requests_[i]) for i in order}
        out = {i: srv.wait(ids[i]) for i in order}
    finally:
        srv.close()
    x2, c2, lat2 = out[7]
    say(f"in a mix of {len(requests_)}: latency {lat2:.2f} ms")
    assert np.array_equal(x1, x2) and np.array_equal(c1, c2)
    b = dev.Batch([probe])
    b.smooth("oddeven")
    xb, cb = b.download(0)
    b.free()
    assert np.array_equal(x1, xb) and np.array_equal(c1, cb)
    oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
    try:
        for i in (0, 7, 13, 29):
            ox, oc = oracle.oddeven(requests_[i])
            x, c, _ = out[i]
            assert max_rel_err(list(x), ox) < TOL and max_rel_err(list(c), oc) < TOL, requests_[i].name
    finally:
        oracle.lib.oracle_set_threads(1)


def test_mixed_precision_server(dev, requests_):
    ref = {}
    b = dev.Batch(requests_[:12])
    b.smooth("oddeven")
    for i in range(12):
        ref[i] = b.download(i, want_cov=False)[0]
    b.free()
    srv = dev.Server(hbm_budget_bytes=2 << 30, max_batch=5, max_inflight=2, mixed_refinements=2)
    try:
        ids = [srv.submit(r) for r in requests_[:12]]
        res = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for i, (x, cv, lat) in enumerate(res):
        e = max_rel_err(list(x), list(ref[i]))
        assert e < TOL, (requests_[i].name, e)
        assert np.all(np.isnan(cv))  # estimates only
    say(f"mixed-precision server: {len(res)} requests within {TOL}")


def test_budget_rejection(dev):
    big = P.serve_problem(dict(n=16, m=8, k=20000, seed=5, unobserved=0.0, name="big"))
    small = P.serve_problem(dict(n=4, m=2, k=50, seed=6, unobserved=0.0, name="small"))
    srv = dev.Server(hbm_budget_bytes=64 << 20, max_batch=4, max_inflight=2)
    try:
        rb, rs = srv.submit(big), srv.submit(small)
        with pytest.raises(dev.UkhipError, match="budget"):
            srv.wait(rb)
        x, cv, lat = srv.wait(rs)
        assert np.all(np.isfinite(x))
    finally:
        srv.close()


def test_bad_request_leaves_others_alone(dev, oracle, requests_):
    """A malformed request (a per-step observation count above its m) is
    refused by ukhip_server_submit itself, so it never joins a batch; the
    valid requests submitted around it all complete and match the oracle."""
    import copy
    srv = dev.Server(hbm_budget_bytes=2 << 30, max_batch=8, max_inflight=2)
    try:
        good = requests_[:6]
        ids = [srv.submit(p) for p in good[:3]]
        bad = copy.copy(good[3])
        bad.packed = dict(good[3].packed)
        mi = np.where(bad.packed["observed"], bad.packed["m"], 0).astype(np.int32)
        mi[len(mi) // 2] = bad.packed["m"] + 1
        bad.packed["m_i"] = mi
        with pytest.raises(dev.UkhipError, match="outside 0..m"):
            srv.submit(bad)
        ids += [srv.submit(p) for p in good[3:]]
        outs = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for p, (x, c, _) in zip(good, outs):
        ox, oc = oracle.oddeven(p)
        assert max_rel_err(list(x), ox) < TOL and max_rel_err(list(c), oc) < TOL, p.name
