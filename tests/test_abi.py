"""The C ABI without a GPU: libukhip loads, exports every function the
public headers declare, and its host-only parts behave like the reference's
(c/matrix_ops.c, c/kalman_base.c). Calls that need the device are expected to
fail loudly through ukhip_last_error when no GPU is present, never to fall
back to a CPU path.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ukhip import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ultimate-kalman_amd", "libukhip.so")
HEADERS = ["kalman.h", "matrix_ops.h", "parallel.h", "ukhip.h"]

KM = C.POINTER(capi.KalmanMatrix)


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#.*", "", text)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name in ("if", "sizeof", "while", "for", "return"):
                continue
            # a declaration: name preceded by a type on the same statement
            start = text.rfind(";", 0, m.start())
            stmt = text[start + 1:m.start()]
            if re.search(r"[A-Za-z_*]\s*$", stmt.strip() + " ") and not stmt.strip().startswith("typedef"):
                names.add(name)
    return names


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    L.ukhip_last_error.restype = C.c_char_p
    for f in ("matrix_create", "matrix_create_constant", "matrix_create_identity"):
        getattr(L, f).restype = KM
        getattr(L, f).argtypes = [C.c_int32, C.c_int32] + ([C.c_double] if f == "matrix_create_constant" else [])
    for f in ("matrix_create_multiply", "matrix_create_add", "matrix_create_subtract", "matrix_create_mldivide",
              "matrix_create_vconcat"):
        getattr(L, f).restype = KM
        getattr(L, f).argtypes = [KM, KM]
    for f in ("matrix_create_inverse", "matrix_create_transpose", "matrix_create_chol", "matrix_create_copy",
              "matrix_create_mutate_qr"):
        getattr(L, f).restype = KM
        getattr(L, f).argtypes = [KM]
    L.matrix_create_from_rowwise.restype = KM
    L.matrix_create_from_rowwise.argtypes = [C.POINTER(C.c_double), C.c_int32, C.c_int32]
    L.matrix_get.restype = C.c_double
    L.matrix_get.argtypes = [KM, C.c_int32, C.c_int32]
    L.matrix_set.argtypes = [KM, C.c_int32, C.c_int32, C.c_double]
    L.matrix_free.argtypes = [KM]
    L.matrix_mutate_apply_qt.argtypes = [KM, KM, KM]
    L.kalman_covariance_matrix_weigh.restype = KM
    L.kalman_covariance_matrix_weigh.argtypes = [KM, C.c_char, KM]
    L.kalman_covariance_matrix_explicit.restype = KM
    L.kalman_covariance_matrix_explicit.argtypes = [KM, C.c_char]
    L.kalman_create_options.restype = C.c_void_p
    L.kalman_create_options.argtypes = [C.c_uint32]
    L.kalman_free.argtypes = [C.c_void_p]
    L.ukhip_get_blocksize.restype = C.c_int
    return L


def to_np(p):
    return capi._read(p)


def from_np(L, a):
    a = np.atleast_2d(np.asarray(a, dtype=np.float64))
    flat = np.ascontiguousarray(a).ravel()
    return L.matrix_create_from_rowwise(flat.ctypes.data_as(C.POINTER(C.c_double)), a.shape[0], a.shape[1])


def test_exports_every_declared_function(lib):
    names = declared_functions()
    assert len(names) >= 60, sorted(names)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert C.c_double.in_dll(lib, "kalman_nan").value != C.c_double.in_dll(lib, "kalman_nan").value  # NaN


def test_matrix_helpers(lib):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((5, 5)) + 5 * np.eye(5)
    B = rng.standard_normal((5, 3))
    a, b = from_np(lib, A), from_np(lib, B)
    assert np.allclose(to_np(a), A) and lib.matrix_get(a, 1, 2) == A[1, 2]
    lib.matrix_set(a, 1, 2, 7.0)
    A[1, 2] = 7.0
    prod = lib.matrix_create_multiply(a, b)
    assert np.allclose(to_np(prod), A @ B, rtol=1e-14, atol=1e-13)
    x = lib.matrix_create_mldivide(a, b)
    assert np.allclose(to_np(x), np.linalg.solve(A, B), rtol=1e-12, atol=1e-12)
    inv = lib.matrix_create_inverse(a)
    assert np.allclose(to_np(inv), np.linalg.inv(A), rtol=1e-12, atol=1e-12)
    S = A @ A.T
    s = from_np(lib, S)
    ch = lib.matrix_create_chol(s)
    Lc = np.tril(to_np(ch))
    assert np.allclose(Lc @ Lc.T, S, rtol=1e-12, atol=1e-12)
    tr = lib.matrix_create_transpose(b)
    assert np.array_equal(to_np(tr), B.T)
    vc = lib.matrix_create_vconcat(a, from_np(lib, B.T))
    assert np.array_equal(to_np(vc), np.vstack([A, B.T]))
    assert not lib.matrix_create_vconcat(None, None)
    for p in (a, b, prod, x, inv, s, ch, tr, vc):
        lib.matrix_free(p)
    lib.matrix_free(None)


def test_qr_apply_qt(lib):
    """matrix_create_mutate_qr + matrix_mutate_apply_qt solve least squares
    like the reference's dgeqrf/dormqr pair (c/matrix_ops.c:382-461)."""
    rng = np.random.default_rng(6)
    A = rng.standard_normal((9, 4))
    y = rng.standard_normal((9, 1))
    qr = from_np(lib, A)
    tau = lib.matrix_create_mutate_qr(qr)
    c = from_np(lib, y)
    lib.matrix_mutate_apply_qt(qr, tau, c)
    R = np.triu(to_np(qr)[:4])
    x = np.linalg.solve(R, to_np(c)[:4])
    assert np.allclose(x, np.linalg.lstsq(A, y, rcond=None)[0], rtol=1e-12, atol=1e-12)
    for p in (qr, tau, c):
        lib.matrix_free(p)


@pytest.mark.parametrize("kind", ["W", "w", "C", "U", "F"])
def test_covariance_weigh_and_explicit(lib, kind):
    """kalman_covariance_matrix_weigh/explicit (c/kalman_base.c:61-214):
    W^T W = Cov^{-1} for the weighing W implied by each representation."""
    rng = np.random.default_rng(7)
    n = 4
    Lt = np.tril(rng.standard_normal((n, n))) + 3 * np.eye(n)
    cov = Lt @ Lt.T
    if kind == "W":
        rep = np.linalg.inv(np.linalg.cholesky(cov))  # any W with W^T W = cov^{-1}
    elif kind == "w":
        cov = np.diag(np.diag(cov))
        rep = np.sqrt(1.0 / np.diag(cov)).reshape(n, 1)
    elif kind == "C":
        rep = cov
    else:  # 'U' and 'F' both hold a lower-triangular L with cov = L L^T (c/kalman_base.c:107-112, :191-199)
        rep = np.linalg.cholesky(cov)
    r = from_np(lib, rep)
    A = rng.standard_normal((n, 3))
    WA = to_np(lib.kalman_covariance_matrix_weigh(r, kind.encode(), from_np(lib, A)))
    # (W A)^T (W A) = A^T cov^{-1} A for every valid W
    assert np.allclose(WA.T @ WA, A.T @ np.linalg.solve(cov, A), rtol=1e-11, atol=1e-11)
    E = to_np(lib.kalman_covariance_matrix_explicit(r, kind.encode()))
    assert np.allclose(E, cov, rtol=1e-11, atol=1e-11)


def test_parallel_hints_and_algorithms(lib):
    lib.parallel_set_blocksize(64)
    assert lib.ukhip_get_blocksize() == 64
    lib.parallel_set_blocksize(16)
    lib.parallel_set_thread_limit(4)
    lib.ukhip_clear_error()
    assert not lib.kalman_create_options(capi.KALMAN_ALGORITHM_ASSOCIATIVE >> 2)  # ULTIMATE (bit 1)
    assert b"ODDEVEN" in lib.ukhip_last_error()
    k = lib.kalman_create_options(capi.KALMAN_ALGORITHM_ASSOCIATIVE)
    assert k
    lib.kalman_free(k)
    # kalman_create() asks for the reference's default, ULTIMATE (c/kalman_base.c:231-232): NULL and an error
    lib.ukhip_clear_error()
    lib.kalman_create.restype = C.c_void_p
    assert not lib.kalman_create()
    assert b"ODDEVEN" in lib.ukhip_last_error()


def test_batch_inputs_checked_before_the_device():
    """ukhip_batch_create refuses a malformed problem (a per-step observation
    count outside 0..m, an unknown covariance type, a missing array) before
    anything is packed, uploaded or launched, so it fails the same way with or
    without a GPU."""
    from ukhip import device
    from ukhip import problems as P
    L = device.lib()
    pr = P.packed_from_lists(P.random_problem(7, 6, 3, 2, K_kind="w", C_kind="w"))
    keep = device._Keep()
    good = device.problem_struct(pr, keep)
    cases = []
    bad = device.problem_struct(pr, keep)
    mi = np.array([2, 2, 3, 2, 2, 2], dtype=np.int32)  # m_i = 3 > m = 2
    keep.arrays.append(mi)
    bad.m_i = mi.ctypes.data
    cases.append((bad, b"m_i = 3 outside 0..m = 2"))
    bad2 = device.problem_struct(pr, keep)
    bad2.C_type = b"x"
    cases.append((bad2, b"unknown covariance type"))
    bad3 = device.problem_struct(pr, keep)
    bad3.G = None
    cases.append((bad3, b"missing input array"))
    bad4 = device.problem_struct(pr, keep)
    bad4.m = 40  # more observation rows than the device reduces for n = 3
    cases.append((bad4, b"outside 0..16"))
    for b, msg in cases:
        arr = (device.ProblemStruct * 2)(good, b)
        L.ukhip_clear_error()
        assert not L.ukhip_batch_create(2, arr)
        assert msg in L.ukhip_last_error(), L.ukhip_last_error()
        assert b"problem 1" in L.ukhip_last_error()


def test_smoother_without_device_fails_loudly(lib):
    """With no HIP device the smoothers record an error and leave the outputs
    NULL: no silent CPU path."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    from ukhip import problems as P
    capi.bind_smoothers(lib, "kalman_smooth_oddeven", "kalman_smooth_associative", "matrix_free")
    pr = P.random_problem(3, 4, 2, 2)
    for fn in (lib.kalman_smooth_oddeven, lib.kalman_smooth_associative):
        lib.ukhip_clear_error()
        eqs = capi.EquationSet(pr)
        fn(0, eqs.ptrs, pr.k)
        assert b"HIP device" in lib.ukhip_last_error()
        assert all(not eqs.structs[i].state for i in range(pr.k))


def test_launch_table_set_and_get():
    """The QR launch table is host state: entries round-trip without a device,
    variants that are not compiled are refused, and set_launch_table parses
    launch_table()'s form."""
    from ukhip import device
    L = device.lib()
    saved = device.launch_table()
    try:
        assert len(saved) == 24
        assert set(saved.values()) <= {"c1w1", "c1w2", "c1w4", "c2w2", "c2w3", "c2w4", "c3w2", "c3w3", "c0w1"}
        cpl = C.c_int(0)
        assert L.ukhip_qr_launch_set(1, 2, 48, 2, 3) == 0
        assert L.ukhip_qr_launch_get(1, 2, 48, C.byref(cpl)) == 3 and cpl.value == 2
        assert L.ukhip_qr_launch_set(0, 2, 48, 1, 3) == -1   # not compiled
        assert L.ukhip_qr_launch_set(0, 2, 48, 2, 1) == -1
        assert L.ukhip_qr_launch_set(0, 2, 48, 3, 2) == 0    # three columns per group
        assert L.ukhip_qr_launch_get(0, 2, 48, C.byref(cpl)) == 2 and cpl.value == 3
        assert L.ukhip_qr_launch_set(0, 2, 48, 3, 1) == -1   # not compiled
        assert L.ukhip_qr_launch_set(0, 2, 48, 0, 2) == -1   # blocked WY: one workgroup per CU only
        assert L.ukhip_qr_launch_set(0, 1, 48, 0, 1) == -1   # and for phase 2 only
        assert L.ukhip_qr_launch_set(0, 2, 24, 0, 1) == -1   # and buckets 32, 48, 64
        assert L.ukhip_qr_launch_set(0, 2, 48, 0, 1) == 0
        assert L.ukhip_qr_launch_get(0, 2, 48, C.byref(cpl)) == 1 and cpl.value == 0
        assert L.ukhip_qr_launch_set(0, 4, 48, 1, 1) == -1
        assert L.ukhip_qr_launch_set(1, 2, 64, 2, 3) == -1   # left out of the FP32 table at bucket 64
        assert L.ukhip_qr_launch_set(0, 2, 64, 2, 3) == 0
        assert L.ukhip_qr_launch_set(1, 1, 64, 2, 3) == 0
        assert L.ukhip_qr_launch_set(0, 1, 8, 1, 1) == -1     # n <= 16 has no variants
        assert L.ukhip_qr_launch_get(0, 1, 8, None) == 0
        device.set_launch_table({"phase3<64, double>": "c2w4", "phase1<32, float>": "c1w1"})
        t = device.launch_table()
        assert t["phase3<64, double>"] == "c2w4" and t["phase1<32, float>"] == "c1w1"
        with pytest.raises(ValueError):
            device.set_launch_table({"phase2<48>": "c1w1"})
        with pytest.raises(ValueError):
            device.set_launch_table({"phase2<48, double>": 2})
        with pytest.raises(RuntimeError):
            device.set_launch_table({"phase2<48, double>": "c1w8"})
    finally:
        device.set_launch_table(saved)
    assert device.launch_table() == saved
