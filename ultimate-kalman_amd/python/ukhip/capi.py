"""ctypes mirror of the kalman.h step-equation interface.

The structs follow c/kalman.h:56-75 and c/matrix_ops.h:137-147 (the same
layout as include/kalman.h with int32 step indices), so one `EquationSet`
can be handed to any library exporting the reference's smoother signature:
libukhip (the product), the CPU oracle and the reference build used to make
golden fixtures.
"""
import ctypes as C

import numpy as np


class KalmanMatrix(C.Structure):
    _fields_ = [("row_dim", C.c_int32), ("col_dim", C.c_int32), ("ld", C.c_int32),
                ("elements", C.POINTER(C.c_double))]


class StepEquations(C.Structure):
    _fields_ = [("step", C.c_int32), ("dimension", C.c_int32),
                ("C_type", C.c_char), ("K_type", C.c_char),
                ("H", C.POINTER(KalmanMatrix)), ("F", C.POINTER(KalmanMatrix)),
                ("K", C.POINTER(KalmanMatrix)), ("c", C.POINTER(KalmanMatrix)),
                ("G", C.POINTER(KalmanMatrix)), ("o", C.POINTER(KalmanMatrix)),
                ("C", C.POINTER(KalmanMatrix)),
                ("state", C.POINTER(KalmanMatrix)), ("covariance", C.POINTER(KalmanMatrix)),
                ("covariance_type", C.c_char)]


SMOOTHER_ARGTYPES = [C.c_uint32, C.POINTER(C.POINTER(StepEquations)), C.c_int32]

KALMAN_ALGORITHM_ODDEVEN = 1 << 2
KALMAN_ALGORITHM_ASSOCIATIVE = 1 << 3
KALMAN_NO_COVARIANCE = 1 << 16
UKHIP_COVARIANCE_W = 1 << 24  # include/ukhip.h: covariances in the 'W' representation


def _as_fortran(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)


class EquationSet:
    """Owns numpy buffers and the kalman_step_equations_t* array built on them.

    `problem` is a ukhip.problems.Problem. Input matrices are borrowed by the
    C side (never freed by it); outputs are allocated by the library that ran
    the smoother and must be released with that library's free function.
    """

    def __init__(self, problem):
        self.problem = problem
        k = problem.k
        self._keep = []
        self._shared = {}  # id(array) -> matrix pointer: a matrix object used on many steps is converted once
        self.structs = (StepEquations * k)()
        self.ptrs = (C.POINTER(StepEquations) * k)()
        for i in range(k):
            s = self.structs[i]
            s.step = i
            s.dimension = problem.n[i]
            s.covariance_type = b"C"
            if i > 0:
                s.H = self._mat(problem.H[i])
                s.F = self._mat(problem.F[i])
                s.c = self._mat(problem.c[i])
                s.K = self._mat(problem.K[i])
                s.K_type = problem.K_type[i].encode()
            if problem.o[i] is not None:
                s.G = self._mat(problem.G[i])
                s.o = self._mat(problem.o[i])
                s.C = self._mat(problem.C[i])
                s.C_type = problem.C_type[i].encode()
            self.ptrs[i] = C.pointer(s)

    def _mat(self, a):
        key = id(a)
        hit = self._shared.get(key)
        if hit is not None and hit[0] is a:
            return hit[1]
        src = a
        a = _as_fortran(a)
        self._keep.append(a)
        m = KalmanMatrix(a.shape[0], a.shape[1], a.shape[0],
                         a.ctypes.data_as(C.POINTER(C.c_double)))
        self._keep.append(m)
        ptr = C.pointer(m)
        self._shared[key] = (src, ptr)
        return ptr

    def collect(self, free_fn=None, want_cov=True, read=True):
        """Reads (read=False: skips reading) and, given `free_fn`, releases
        every step's state/covariance."""
        states, covs = [], []
        for i in range(self.problem.k):
            s = self.structs[i]
            x = _read(s.state) if read else None
            states.append(None if x is None else x[:, 0])
            covs.append(_read(s.covariance) if (want_cov and read) else None)
            if free_fn is not None:
                if s.state:
                    free_fn(s.state)
                if s.covariance:
                    free_fn(s.covariance)
            s.state = C.POINTER(KalmanMatrix)()
            s.covariance = C.POINTER(KalmanMatrix)()
        return states, covs


def _read(p):
    if not p:
        return None
    m = p.contents
    if m.row_dim * m.col_dim == 0:
        return np.zeros((m.row_dim, m.col_dim))
    buf = np.ctypeslib.as_array(m.elements, shape=(m.ld * m.col_dim,))
    a = buf.reshape((m.col_dim, m.ld)).T[: m.row_dim, :]
    return np.array(a, copy=True)


def bind_smoothers(lib, oddeven_name, assoc_name, free_name):
    """Returns (oddeven, associative, free) callables from a loaded CDLL."""
    oe = getattr(lib, oddeven_name)
    oe.argtypes = SMOOTHER_ARGTYPES
    oe.restype = None
    asc = getattr(lib, assoc_name)
    asc.argtypes = SMOOTHER_ARGTYPES
    asc.restype = None
    fr = getattr(lib, free_name)
    fr.argtypes = [C.POINTER(KalmanMatrix)]
    fr.restype = None
    return oe, asc, fr


def run(fn, free_fn, problem, options=0):
    """Runs one smoother over `problem`; returns (states, covariances)."""
    eqs = EquationSet(problem)
    fn(options, eqs.ptrs, problem.k)
    return eqs.collect(free_fn, want_cov=not (options & KALMAN_NO_COVARIANCE))
