"""The QR launch table and its run-time autotuner (ukhip_autotune).

Every launch variant (one column per lane group at 1, 2 or 4 workgroups per
CU asked of the register allocator, two columns per group at 2, 3 or 4, three
at 2 or 3) runs
the same operations on every column: results must be bit-identical across
variants. The autotuner must leave a valid table.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from ukhip import problems as P

pytestmark = pytest.mark.gpu

VARIANTS = [(1, 1), (1, 2), (1, 4), (2, 2), (2, 3), (2, 4), (3, 2), (3, 3)]  # (columns per group, workgroups per CU)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    device.lib()
    return device


def _get(L, fp32, p, n):
    cpl = C.c_int(0)
    wg = L.ukhip_qr_launch_get(fp32, p, n, C.byref(cpl))
    return cpl.value, wg


@pytest.mark.parametrize("n,m", [(24, 12), (40, 20), (64, 32)])
def test_variants_bit_identical(dev, n, m):
    L = dev.lib()
    F = 0.97 * P.random_orthogonal(P.rng_for(800 + n), n)
    probs = [P.iid_problem(k, n, m, seed=810 + k, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1,
                           unobserved=lambda i: i % 10 == 9, name=f"n{n}_k{k}") for k in (3, 200)]
    saved = {p: _get(L, 0, p, n) for p in (1, 2, 3)}
    L.ukhip_use_graphs(0)
    try:
        ref = None
        for p in (1, 2, 3):
            for cpl, wg in VARIANTS:
                assert L.ukhip_qr_launch_set(0, p, n, cpl, wg) == 0
                b = dev.Batch(probs)
                b.smooth("oddeven")
                out = [b.download(i) for i in range(len(probs))]
                b.free()
                L.ukhip_qr_launch_set(0, p, n, *saved[p])
                if ref is None:
                    ref = out
                for (x0, c0), (x1, c1) in zip(ref, out):
                    assert np.array_equal(x0, x1) and np.array_equal(c0, c1), (p, cpl, wg)
        say(f"n={n}: {3 * len(VARIANTS)} launch variants bit-identical")
    finally:
        for p, v in saved.items():
            L.ukhip_qr_launch_set(0, p, n, *v)
        L.ukhip_use_graphs(1)


@pytest.mark.parametrize("n,m,illcond", [(32, 16, False), (40, 20, False), (48, 24, False), (64, 32, True)])
def test_fp32_variants_bit_identical(dev, n, m, illcond):
    """The FP32 elimination's launch variants: identical, finite mixed-precision
    estimates (n = 64: the ill-conditioned dynamics F = Q diag(logspace(0, -3,
    64)), K = 1e-2 I). Variants the launch table refuses are skipped."""
    L = dev.lib()
    if illcond:
        F = P.random_orthogonal(P.rng_for(25), n) @ np.diag(np.logspace(0, -3, n))
        pr = P.iid_problem(257, n, m, seed=26, F=F, G_scale=1 / 8.0, Kvar=1e-2, name=f"n{n}_illcond_k257")
    else:
        F = 0.97 * P.random_orthogonal(P.rng_for(820 + n), n)
        pr = P.iid_problem(301, n, m, seed=830, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1, name=f"n{n}_k301")
    saved = {p: _get(L, 1, p, n) for p in (1, 2, 3)}
    try:
        ref = None
        for p in (1, 2, 3):
            for cpl, wg in VARIANTS:
                if L.ukhip_qr_launch_set(1, p, n, cpl, wg) != 0:
                    say(f"  phase {p} c{cpl}w{wg}: not in the launch table")
                    continue
                b = dev.Batch([pr])
                b.smooth_mixed(1)
                x = b.download(0, want_cov=False)[0]
                b.free()
                L.ukhip_qr_launch_set(1, p, n, *saved[p])
                assert np.isfinite(x).all(), (p, cpl, wg)
                if ref is None:
                    ref = x
                assert np.array_equal(ref, x), (p, cpl, wg)
        say(f"n={n}: FP32 launch variants bit-identical")
    finally:
        for p, v in saved.items():
            L.ukhip_qr_launch_set(1, p, n, *v)


def test_autotune_table(dev):
    table = dev.autotune(64, 1024)
    say(f"autotuned launch table: {table}")
    assert set(table.values()) <= {f"c{c}w{w}" for c, w in VARIANTS} | {"c0w1"}
    assert all(v != "c0w1" for k, v in table.items() if not k.startswith("phase2")), table
    assert table["phase2<64, float>"] != "c2w3", table
