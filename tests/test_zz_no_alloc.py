"""The allocation-free entry point ukhip_smooth_oddeven_noalloc (include/ukhip.h)
against the batch path it is a variant of (ukhip_batch_create +
ukhip_batch_smooth_oddeven + ukhip_batch_download) on the same problem.

Both run the same kernels on the same layout, so the estimates and
covariances must agree bit for bit. Two calls in a row, the second on a
smaller problem (the scratch keeps the first call's contents), and a problem
larger than the scratch is refused with an error before anything is launched.
Inputs and outputs live in device memory allocated by the test itself.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from ukhip import problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime (the test's
    own device buffers)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.restype = C.c_int
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.live.append(p)
        return p.value

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        self.live.append(p)
        return p.value

    def get(self, ptr, shape):
        out = np.empty(shape)
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


@pytest.fixture(scope="module")
def dev():
    from ukhip import device
    L = device.lib()
    L.ukhip_smooth_oddeven_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
    L.ukhip_smooth_oddeven_noalloc.restype = C.c_int
    return device


def _problem(k, n, m, seed, kind):
    F = 0.98 * P.random_orthogonal(P.rng_for(seed), n)
    return P.iid_problem(k, n, m, seed=seed + 1, F=F, G_scale=1 / np.sqrt(n), Kvar=1.0, Cvar=0.1, K_kind=kind,
                         C_kind=kind, unobserved=lambda i: i % 7 == 3)


def _device_struct(dev, hip, pr):
    """ukhip_problem_t of pr with every array copied to device memory."""
    keep = dev._Keep()
    s = dev.problem_struct(pr, keep)
    k, n, m = pr.k, pr.packed["n"], pr.packed["m"]
    kw = n if pr.packed["K_type"] == "w" else n * n
    cw = m if pr.packed["C_type"] == "w" else m * m
    sizes = {"F": (n * n if s.F_shared else k * n * n) * 8, "c": k * n * 8, "K": k * kw * 8, "G": k * m * n * 8,
             "o": k * m * 8, "C": k * cw * 8, "m_i": k * 4}
    for f, nbytes in sizes.items():
        host = getattr(s, f)
        buf = (C.c_char * nbytes).from_address(host)
        setattr(s, f, hip.put(np.frombuffer(bytes(buf), dtype=np.uint8)))
    return s, keep


def _batch_result(dev, pr, cov):
    b = dev.Batch([pr])
    try:
        b.smooth("oddeven", options=0 if cov else (1 << 16))
        return b.download(0, want_cov=cov)
    finally:
        b.free()


def _noalloc_result(dev, hip, pr, cov):
    s, _keep = _device_struct(dev, hip, pr)
    k, n = pr.k, pr.packed["n"]
    xs = hip.empty(k * n * 8)
    Ss = hip.empty(k * n * n * 8) if cov else None
    rc = dev.lib().ukhip_smooth_oddeven_noalloc(C.addressof(s), 0 if cov else (1 << 16), xs, Ss, None)
    assert rc == 0, dev.last_error()
    assert dev.lib().ukhip_synchronize() == 0
    x = hip.get(xs, (k, n))
    S = np.swapaxes(hip.get(Ss, (k, n, n)), 1, 2) if cov else None
    return x, S


@pytest.mark.parametrize("kind", ["w", "C"])
def test_noalloc_matches_batch(dev, kind):
    hip = Hip()
    try:
        for (k, n, m, seed) in ((1500, 12, 5, 31), (77, 6, 6, 41)):  # the second call is smaller
            pr = _problem(k, n, m, seed, kind)
            for cov in (True, False):
                say(f"noalloc n={n} m={m} k={k} kind={kind} cov={cov}")
                xa, Sa = _noalloc_result(dev, hip, pr, cov)
                xb, Sb = _batch_result(dev, pr, cov)
                assert np.array_equal(xa, xb), (k, n, cov)
                if cov:
                    assert np.array_equal(Sa, Sb), (k, n)
                assert np.all(np.isfinite(xa))
    finally:
        hip.free()


def test_noalloc_refuses_beyond_scratch(dev):
    hip = Hip()
    try:
        pr = _problem(4096, 48, 24, 51, "w")  # about 0.7 GB of intermediates: beyond the scratch
        s, _keep = _device_struct(dev, hip, pr)
        xs = hip.empty(pr.k * 48 * 8)
        dev.lib().ukhip_clear_error()
        rc = dev.lib().ukhip_smooth_oddeven_noalloc(C.addressof(s), 1 << 16, xs, None, None)
        assert rc == -1
        assert "bytes of device scratch" in dev.last_error()
    finally:
        hip.free()
