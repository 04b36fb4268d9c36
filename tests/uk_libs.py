"""Loaders for the three libraries the tests drive through the same C ABI.

- product:   ultimate-kalman_amd/libukhip.so (HIP; needs a GPU to compute)
- oracle:    oracle/liboracle.so (CPU FP64 restatement; the checker)
- reference: oracle/_ref/libukref.so (the reference C, built from its own
             sources in this container; absent on machines without them)
"""
import ctypes as C
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ultimate-kalman_amd")
sys.path.insert(0, os.path.join(PKG, "python"))

from ukhip import capi  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libukref.so")
PRODUCT_SO = os.path.join(PKG, "libukhip.so")

_cache = {}


class Smoothers:
    def __init__(self, lib, oe, asc, free):
        self.lib = lib
        self.oe, self.asc, self.free = capi.bind_smoothers(lib, oe, asc, free)

    def oddeven(self, problem, options=0):
        return capi.run(self.oe, self.free, problem, options)

    def associative(self, problem, options=0):
        return capi.run(self.asc, self.free, problem, options)


def oracle():
    if "oracle" not in _cache:
        lib = C.CDLL(ORACLE_SO)
        s = Smoothers(lib, "oracle_smooth_oddeven", "oracle_smooth_associative", "oracle_matrix_free")
        lib.oracle_count_flops.argtypes = [C.c_int]
        lib.oracle_flops.restype = C.c_double
        lib.oracle_set_threads.argtypes = [C.c_int]
        lib.oracle_dense_ls.argtypes = [C.POINTER(C.POINTER(capi.StepEquations)), C.c_int32,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.oracle_dense_ls.restype = C.c_int
        _cache["oracle"] = s
    return _cache["oracle"]


def reference():
    if "ref" not in _cache:
        if not os.path.exists(REF_SO):
            return None
        lib = C.CDLL(REF_SO)
        _cache["ref"] = Smoothers(lib, "kalman_smooth_oddeven", "kalman_smooth_associative", "matrix_free")
    return _cache["ref"]


def dense_ls(problem):
    """Global least-squares states and covariance blocks (oracle_dense_ls)."""
    import numpy as np
    o = oracle()
    eqs = capi.EquationSet(problem)
    N = sum(problem.n)
    st = np.zeros(N)
    cv = np.zeros(sum(n * n for n in problem.n))
    rc = o.lib.oracle_dense_ls(eqs.ptrs, problem.k, st.ctypes.data_as(C.POINTER(C.c_double)),
                               cv.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("oracle_dense_ls failed")
    states, covs, a, b = [], [], 0, 0
    for n in problem.n:
        states.append(st[a:a + n])
        covs.append(cv[b:b + n * n].reshape(n, n).T)
        a += n
        b += n * n
    return states, covs
