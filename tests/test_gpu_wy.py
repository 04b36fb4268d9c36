"""Phase 2 of the elimination in blocked WY form on the matrix cores
(k_phase2_wy, launch-table variant c0w1) against the register/DPP kernel it
replaces and the oracle.

Bars: FP64 estimates and covariances within 1e-9 of the oracle and 1e-12 of
the register kernel (the same reflectors applied in another order); the FP32
elimination in WY form, after two FP64 refinement steps, within 1e-9 of the
FP64 estimates. Buckets 32, 48 and 64 (bucket 24 has no blocked form);
diagonal whitening (level-0 D stored as its diagonal, B formed on load) and
dense whitening; odd lengths (lone last steps) and short problems.
"""
import ctypes as C
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


def _set_phase2(dev, fp32, n, variant):
    """Sets the phase-2 launch-table entry; returns the previous one."""
    L = dev.lib()
    cpl = C.c_int(0)
    wg = L.ukhip_qr_launch_get(fp32, 2, n, C.byref(cpl))
    assert L.ukhip_qr_launch_set(fp32, 2, n, variant[0], variant[1]) == 0
    return (cpl.value, wg)


@pytest.mark.parametrize("n,m,kind", [(28, 14, "w"), (32, 16, "C"), (40, 20, "w"), (48, 24, "C"), (64, 32, "w")])
def test_wy_phase2_matches(dev, oracle, n, m, kind):
    F = 0.98 * P.random_orthogonal(P.rng_for(900 + n), n)
    probs = [P.iid_problem(k, n, m, seed=910 + k + n, F=F, G_scale=1 / np.sqrt(n), Cvar=0.1, K_kind=kind,
                           C_kind=kind, unobserved=lambda i: i % 10 == 9, name=f"n{n}_{kind}_k{k}")
             for k in (2, 3, 7, 130, 333)]
    L = dev.lib()
    L.ukhip_use_graphs(0)  # launches pick up the launch table as it is now
    try:
        b = dev.Batch(probs)
        ref = []
        b.smooth("oddeven")
        for i in range(len(probs)):
            ref.append(b.download(i))
        old64 = _set_phase2(dev, 0, n, (0, 1))
        old32 = _set_phase2(dev, 1, n, (0, 1))
        try:
            b.smooth("oddeven")
            wy = [b.download(i) for i in range(len(probs))]
            b.smooth_mixed_opt(2, 0, False)
            mx = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
        finally:
            _set_phase2(dev, 0, n, old64)
            _set_phase2(dev, 1, n, old32)
        b.free()
        oracle.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
        for i, pr in enumerate(probs):
            (x, S), (xr, Sr) = wy[i], ref[i]
            d = max(max_rel_err(list(x), list(xr)), max_rel_err(list(S), list(Sr)))
            ox, oc = oracle.oddeven(pr)
            eo = max(max_rel_err(list(x), ox), max_rel_err(list(S), oc))
            em = max_rel_err(list(mx[i]), list(xr))
            say(f"  {pr.name}: WY vs register kernel {d:.1e}, vs oracle {eo:.1e}; FP32 WY + 2 refinements {em:.1e}")
            assert d < 1e-12 and eo < TOL and em < TOL, pr.name
    finally:
        oracle.lib.oracle_set_threads(1)
        L.ukhip_use_graphs(1)
