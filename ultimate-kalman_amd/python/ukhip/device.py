"""ctypes binding of libukhip (include/kalman.h, include/ukhip.h).

Loads the in-tree ultimate-kalman_amd/libukhip.so. There is no fallback: if
the library or a HIP device is missing, the calls fail loudly.
"""
import ctypes as C
import ctypes as _ctypes  # for methods whose arguments shadow C
import os

import numpy as np

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(os.path.dirname(HERE))
LIB_PATH = os.path.join(PKG, "libukhip.so")


class ProblemStruct(C.Structure):
    """ukhip_problem_t (include/ukhip.h)."""
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("k", C.c_int64),
                ("F", C.c_void_p), ("H", C.c_void_p), ("c", C.c_void_p), ("K", C.c_void_p),
                ("K_type", C.c_char), ("G", C.c_void_p), ("o", C.c_void_p), ("m_i", C.c_void_p),
                ("C", C.c_void_p), ("C_type", C.c_char), ("F_shared", C.c_int32)]


_lib = None

# slots of ukhip_batch_smooth_oddeven_timed (include/ukhip.h)
OE_KERNELS = ("k_weigh", "k_phase1", "k_phase2", "k_phase3", "k_backsub", "total", "k_solve_est", "k_normal",
              "k_fwd", "convert")  # the UKHIP_OE_* slots of include/ukhip.h, in order
# slots of ukhip_batch_smooth_associative_timed (UKHIP_AS_*), in order
AS_KERNELS = ("weigh", "model", "prior", "felem", "fscan", "filtered", "selem", "sscan", "results", "total")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.ukhip_last_error.restype = C.c_char_p
        L.ukhip_batch_create.restype = C.c_void_p
        L.ukhip_batch_create.argtypes = [C.c_int32, C.POINTER(ProblemStruct)]
        L.ukhip_batch_free.argtypes = [C.c_void_p]
        L.ukhip_batch_steps.restype = C.c_int64
        L.ukhip_batch_steps.argtypes = [C.c_void_p]
        L.ukhip_batch_levels.restype = C.c_int32
        L.ukhip_batch_levels.argtypes = [C.c_void_p]
        for f in ("ukhip_batch_smooth_oddeven", "ukhip_batch_smooth_associative"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_uint32]
            getattr(L, f).restype = C.c_int
        L.ukhip_batch_smooth_oddeven_timed.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_double),
                                                       C.POINTER(C.c_int32)]
        L.ukhip_batch_smooth_oddeven_timed.restype = C.c_int
        L.ukhip_batch_timed_parts.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.POINTER(C.c_double),
                                              C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32]
        L.ukhip_batch_timed_parts.restype = C.c_int
        L.ukhip_batch_smooth_associative_timed.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_double),
                                                           C.POINTER(C.c_int32)]
        L.ukhip_batch_smooth_associative_timed.restype = C.c_int
        L.ukhip_batch_download.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.ukhip_batch_download.restype = C.c_int
        L.ukhip_batch_download_cov.argtypes = [C.c_void_p, C.c_int32, C.c_char, C.c_void_p]
        L.ukhip_batch_download_cov.restype = C.c_int
        L.ukhip_synchronize.restype = C.c_int
        L.ukhip_stream.restype = C.c_void_p
        L.ukhip_use_graphs.argtypes = [C.c_int]
        L.ukhip_use_mfma.argtypes = [C.c_int]
        L.ukhip_batch_smooth_oddeven_mixed.argtypes = [C.c_void_p, C.c_int32]
        L.ukhip_batch_smooth_oddeven_mixed.restype = C.c_int
        L.ukhip_batch_smooth_oddeven_mixed_opt.argtypes = [C.c_void_p, C.c_void_p]
        L.ukhip_batch_smooth_oddeven_mixed_opt.restype = C.c_int
        L.ukhip_server_create.restype = C.c_void_p
        L.ukhip_server_create.argtypes = [C.c_void_p]
        L.ukhip_server_submit.restype = C.c_int64
        L.ukhip_server_submit.argtypes = [C.c_void_p, C.POINTER(ProblemStruct)]
        L.ukhip_server_poll.argtypes = [C.c_void_p, C.c_int64]
        L.ukhip_server_wait.restype = C.c_int
        L.ukhip_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.ukhip_server_destroy.argtypes = [C.c_void_p]
        L.ukhip_autotune.argtypes = [C.c_int32, C.c_int32]
        L.ukhip_autotune.restype = C.c_int
        L.ukhip_qr_launch_get.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.ukhip_qr_launch_set.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ukhip_filter_create.restype = C.c_void_p
        L.ukhip_filter_create.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        L.ukhip_filter_step.argtypes = [C.c_void_p, C.c_void_p]
        L.ukhip_filter_step.restype = C.c_int
        L.ukhip_filter_step_device.argtypes = [C.c_void_p, C.c_void_p]
        L.ukhip_filter_step_device.restype = C.c_int
        L.ukhip_filter_estimates.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ukhip_filter_estimates.restype = C.c_int
        L.ukhip_filter_free.argtypes = [C.c_void_p]
        L.ukhip_filter_keep_history.argtypes = [C.c_void_p, C.c_int64]
        L.ukhip_filter_keep_history.restype = C.c_int
        L.ukhip_filter_smooth.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ukhip_filter_smooth.restype = C.c_int
        capi.bind_smoothers(L, "kalman_smooth_oddeven", "kalman_smooth_associative", "matrix_free")
        _lib = L
    return _lib


def autotune(n_max=64, k=4096):
    """ukhip_autotune; returns the launch table (launch_table())."""
    L = lib()
    _check(L.ukhip_autotune(n_max, k), "ukhip_autotune")
    return launch_table()


def launch_table():
    """The QR launch table: {"phase<p><<bucket>, double|float>": "c<columns per group>w<workgroups per CU>"}."""
    L = lib()
    out = {}
    for f, t in ((0, "double"), (1, "float")):
        for nb in (24, 32, 48, 64):
            for p in (1, 2, 3):
                cpl = C.c_int(0)
                wg = L.ukhip_qr_launch_get(f, p, nb, C.byref(cpl))
                out[f"phase{p}<{nb}, {t}>"] = f"c{cpl.value}w{wg}"
    return out


def set_launch_table(table):
    """Sets launch-table entries from a mapping in launch_table()'s form
    ({"phase2<48, double>": "c2w3", ...}), e.g. one saved from an earlier autotune."""
    import re
    L = lib()
    for key, val in table.items():
        mt = re.fullmatch(r"phase([123])<(\d+), (double|float)>", key.strip())
        mv = re.fullmatch(r"c(\d)w(\d)", str(val).strip())
        if mt is None or mv is None:
            raise ValueError(f"launch table entry {key!r}: {val!r}: expected 'phase<1..3><<bucket>, double|float>': "
                             "'c<columns per group>w<workgroups per CU>'")
        _check(L.ukhip_qr_launch_set(int(mt.group(3) == "float"), int(mt.group(1)), int(mt.group(2)),
                                     int(mv.group(1)), int(mv.group(2))),
               f"ukhip_qr_launch_set {key}={val}")


def last_error():
    return lib().ukhip_last_error().decode()


class UkhipError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise UkhipError(f"{what}: {last_error()}")


def smooth_equations(problem, algorithm="oddeven", options=0):
    """The reference's entry point on the GPU: kalman_smooth_<algorithm> over
    kalman_step_equations_t built from `problem`. Returns (states, covs)."""
    L = lib()
    L.ukhip_clear_error()
    fn = L.kalman_smooth_oddeven if algorithm == "oddeven" else L.kalman_smooth_associative
    eqs = capi.EquationSet(problem)
    fn(options, eqs.ptrs, problem.k)
    err = last_error()
    states, covs = eqs.collect(L.matrix_free, want_cov=not (options & capi.KALMAN_NO_COVARIANCE))
    if err:
        raise UkhipError(err)
    return states, covs


class _Keep:
    """Holds the numpy buffers a ProblemStruct points into."""

    def __init__(self):
        self.arrays = []

    def ptr(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.arrays.append(a)
        return a.ctypes.data


def problem_struct(p, keep):
    """ukhip_problem_t of a packed Problem (buffers kept alive by `keep`)."""
    pk = p.packed
    n, m, k = pk["n"], pk["m"], p.k
    s = ProblemStruct()
    s.n, s.m, s.k = n, m, k
    if pk.get("F_shared") is not None:
        s.F = keep.ptr(np.asfortranarray(pk["F_shared"]).T.ravel())  # column-major n x n
        s.F_shared = 1
    else:
        s.F = keep.ptr(pk["F"])
    s.H = keep.ptr(pk.get("H"))
    s.c = keep.ptr(pk["c"])
    s.K_type = pk["K_type"].encode()
    if "K_steps" in pk:  # one K per step (packed_from_lists)
        s.K = keep.ptr(pk["K_steps"])
    else:
        K = pk["K"]
        s.K = keep.ptr(np.broadcast_to(np.asfortranarray(K).T.ravel(), (k, K.size)))
    s.G = keep.ptr(pk["G"])
    s.o = keep.ptr(pk["o"])
    mi = pk["m_i"] if "m_i" in pk else np.where(pk["observed"], m, 0).astype(np.int32)
    mi = np.ascontiguousarray(mi, dtype=np.int32)
    keep.arrays.append(mi)
    s.m_i = mi.ctypes.data
    s.C_type = pk["C_type"].encode()
    if "C_steps" in pk:
        s.C = keep.ptr(pk["C_steps"])
    else:
        Cm = pk["C"]
        s.C = keep.ptr(np.broadcast_to(np.asfortranarray(Cm).T.ravel(), (k, Cm.size)))
    return s


_hip = None


def _hip_lib():
    """The HIP runtime (device memory for the device-resident entry points)."""
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A host array copied once into device memory (hipMalloc + hipMemcpy),
    for the entry points that read device-resident inputs."""

    def __init__(self, a, dtype=np.float64):
        lib()
        a = np.ascontiguousarray(a, dtype=dtype)
        h = _hip_lib()
        p = C.c_void_p()
        if h.hipMalloc(C.byref(p), max(a.nbytes, 1)) != 0:
            raise UkhipError(f"hipMalloc of {a.nbytes} bytes failed")
        self.ptr = p.value
        if h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:  # hipMemcpyHostToDevice
            h.hipFree(self.ptr)
            self.ptr = None
            raise UkhipError("hipMemcpy to the device failed")

    def free(self):
        if self.ptr:
            _hip_lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FilterInputs(C.Structure):
    """ukhip_filter_inputs_t (include/ukhip.h)."""
    _fields_ = [("F", C.c_void_p), ("F_shared", C.c_int32), ("H", C.c_void_p), ("c", C.c_void_p),
                ("K", C.c_void_p), ("G", C.c_void_p), ("o", C.c_void_p), ("m_i", C.c_void_p), ("C", C.c_void_p),
                ("K_type", C.c_char), ("C_type", C.c_char)]


class Filter:
    """Batched sequential UltimateKalman filter (ukhip_filter_*): nfilters
    filters of state dimension n, m observation rows per step at most."""

    def __init__(self, nfilters, n, m):
        L = lib()
        L.ukhip_clear_error()
        self.h = L.ukhip_filter_create(nfilters, n, m)
        if not self.h:
            raise UkhipError("ukhip_filter_create: " + last_error())
        self.nf, self.n, self.m = nfilters, n, m
        self._steps = 0

    def keep_history(self, max_steps):
        """ukhip_filter_keep_history: keep every step's factors for smooth()."""
        _check(lib().ukhip_filter_keep_history(self.h, int(max_steps)), "ukhip_filter_keep_history")

    def smooth(self, want_W=True):
        """ukhip_filter_smooth: (states (steps, nf, n), W (steps, nf, n, n) with
        W^T W the inverse covariance) over the steps so far."""
        steps = self._steps
        x = np.empty((steps, self.nf, self.n))
        W = np.empty((steps, self.nf, self.n, self.n)) if want_W else None
        _check(lib().ukhip_filter_smooth(self.h, x.ctypes.data, W.ctypes.data if want_W else None),
               "ukhip_filter_smooth")
        return x, (np.swapaxes(W, 2, 3) if want_W else None)

    @staticmethod
    def layout(K, G, o, C, F=None, H=None, c=None, m_i=None, K_type="w", C_type="w"):
        """One step's inputs in the C layout of ukhip_filter_inputs_t: a dict
        of contiguous arrays (matrices column-major, filter after filter) plus
        F_shared and the representation codes. Arrays hold all filters: F
        (nf, n, n) or (n, n) shared, H/c optional, G (nf, m, n), o (nf, m),
        m_i (nf,) observation counts; K and C in the representation K_type /
        C_type: 'w' inverse standard deviations (K (nf, n), C (nf, m)), or 'W',
        'C', 'U', 'F' matrices (K (nf, n, n), C (nf, m, m))."""

        def arr(a, dtype=np.float64):
            return None if a is None else np.ascontiguousarray(a, dtype=dtype)

        def colmajor(a):  # (..., rows, cols) -> column-major per matrix
            return None if a is None else arr(np.swapaxes(np.asarray(a, dtype=np.float64), -1, -2))

        Fa = None if F is None else np.asarray(F, dtype=np.float64)
        return {"F": colmajor(Fa), "F_shared": 1 if (Fa is not None and Fa.ndim == 2) else 0, "H": colmajor(H),
                "c": arr(c), "K": arr(K) if K_type == "w" else colmajor(K), "G": colmajor(G), "o": arr(o),
                "C": arr(C) if C_type == "w" else colmajor(C), "m_i": arr(m_i, np.int32),
                "K_type": K_type, "C_type": C_type}

    @staticmethod
    def _struct(lay, ptr):
        s = FilterInputs()
        for k in ("F", "H", "c", "K", "G", "o", "C", "m_i"):
            setattr(s, k, ptr(lay[k]))
        s.F_shared = lay["F_shared"]
        s.K_type = lay["K_type"].encode()
        s.C_type = lay["C_type"].encode()
        return s

    def step(self, K, G, o, C, F=None, H=None, c=None, m_i=None, K_type="w", C_type="w"):
        """One evolve + observe step from host arrays (layout() gives the
        shapes)."""
        lay = self.layout(K, G, o, C, F, H, c, m_i, K_type, C_type)
        s = self._struct(lay, lambda a: None if a is None else a.ctypes.data)
        _check(lib().ukhip_filter_step(self.h, _ctypes.byref(s)), "ukhip_filter_step")
        self._steps += 1

    @staticmethod
    def upload(**kw):
        """One step's inputs (the arguments of step()) copied to device memory
        once: a dict of DeviceArray for step_device."""
        lay = Filter.layout(**kw)
        return {k: (DeviceArray(v, v.dtype) if isinstance(v, np.ndarray) else v) for k, v in lay.items()}

    def step_device(self, dev_inputs):
        """ukhip_filter_step_device: one step from inputs already in device
        memory (upload())."""
        s = self._struct(dev_inputs, lambda a: None if a is None else a.ptr)
        _check(lib().ukhip_filter_step_device(self.h, _ctypes.byref(s)), "ukhip_filter_step_device")
        self._steps += 1

    def estimates(self, want_R=True):
        """(states (nf, n), R (nf, n, n) upper triangular with R^T R the inverse covariance)."""
        x = np.empty((self.nf, self.n))
        R = np.empty((self.nf, self.n, self.n)) if want_R else None
        _check(lib().ukhip_filter_estimates(self.h, x.ctypes.data, R.ctypes.data if want_R else None),
               "ukhip_filter_estimates")
        return x, (np.swapaxes(R, 1, 2) if want_R else None)

    def close(self):
        if self.h:
            lib().ukhip_filter_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MixedOptions(C.Structure):
    """ukhip_mixed_options_t (include/ukhip.h)."""
    _fields_ = [("refinements", C.c_int32), ("fp32_levels", C.c_int32), ("covariances", C.c_int32)]


class ServerOptions(C.Structure):
    """ukhip_server_options_t (include/ukhip.h)."""
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_batch", C.c_int32), ("max_inflight", C.c_int32),
                ("mixed_refinements", C.c_int32)]


class Server:
    """Continuous-batching smoothing server (ukhip_server_*)."""

    def __init__(self, hbm_budget_bytes=0, max_batch=256, max_inflight=4, mixed_refinements=-1):
        L = lib()
        L.ukhip_clear_error()
        o = ServerOptions(hbm_budget_bytes, max_batch, max_inflight, mixed_refinements)
        self.h = L.ukhip_server_create(C.byref(o))
        if not self.h:
            raise UkhipError("ukhip_server_create: " + last_error())
        self.mixed = mixed_refinements >= 0
        self._shapes = {}

    def submit(self, problem):
        keep = _Keep()
        s = problem_struct(problem, keep)
        rid = lib().ukhip_server_submit(self.h, C.byref(s))  # inputs are copied by the call
        if rid < 0:
            raise UkhipError("ukhip_server_submit: " + last_error())
        self._shapes[rid] = (problem.k, problem.packed["n"])
        return rid

    def wait(self, rid, want_cov=True, want_states=True):
        """(states k x n or None, covariances k x n x n or None, latency in ms)."""
        k, n = self._shapes.pop(rid)
        x = np.empty((k, n)) if want_states else None
        cv = np.empty((k, n, n)) if want_cov else None
        lat = C.c_double()
        rc = lib().ukhip_server_wait(self.h, rid, x.ctypes.data if want_states else None,
                                     cv.ctypes.data if want_cov else None, C.byref(lat))
        if rc != 0:
            raise UkhipError(last_error())
        if want_cov:
            cv = np.swapaxes(cv, 1, 2)
        return x, cv, lat.value

    def close(self):
        if self.h:
            lib().ukhip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Device-resident set of problems with one state dimension."""

    def __init__(self, problems):
        L = lib()
        L.ukhip_clear_error()
        self._keep = []
        arr = (ProblemStruct * len(problems))()
        for i, p in enumerate(problems):
            arr[i] = self._struct(p)
        self.problems = problems
        self.h = L.ukhip_batch_create(len(problems), arr)
        if not self.h:
            raise UkhipError("ukhip_batch_create: " + last_error())

    def _struct(self, p):
        if not hasattr(self, "_kp"):
            self._kp = _Keep()
            self._keep.append(self._kp)
        return problem_struct(p, self._kp)

    @property
    def steps(self):
        return lib().ukhip_batch_steps(self.h)

    @property
    def levels(self):
        return lib().ukhip_batch_levels(self.h)

    def smooth(self, algorithm="oddeven", options=0, sync=True):
        L = lib()
        fn = L.ukhip_batch_smooth_oddeven if algorithm == "oddeven" else L.ukhip_batch_smooth_associative
        _check(fn(self.h, options), f"smooth_{algorithm}")
        if sync:
            _check(L.ukhip_synchronize(), "synchronize")

    def smooth_mixed(self, refinements=3, sync=True):
        """Estimates by FP32 elimination plus `refinements` FP64 refinement
        steps (covariances untouched)."""
        L = lib()
        _check(L.ukhip_batch_smooth_oddeven_mixed(self.h, refinements), "smooth_mixed")
        if sync:
            _check(L.ukhip_synchronize(), "synchronize")

    def smooth_mixed_opt(self, refinements=1, fp32_levels=0, covariances=False, sync=True):
        """ukhip_batch_smooth_oddeven_mixed_opt: levels 0..fp32_levels-1 in FP32
        and the rest in FP64 (<= 0: all FP32); with covariances, the covariance
        output is formed from the block factors too."""
        L = lib()
        o = MixedOptions(refinements, fp32_levels, 1 if covariances else 0)
        _check(L.ukhip_batch_smooth_oddeven_mixed_opt(self.h, C.byref(o)), "smooth_mixed_opt")
        if sync:
            _check(L.ukhip_synchronize(), "synchronize")

    def smooth_timed(self, options=0):
        """Odd-even smoothing timed with HIP events around every launch.
        Returns ({kernel: ms}, {kernel: launches}) for k_weigh, k_phase1,
        k_phase2, k_phase3, k_backsub and the whole span ("total")."""
        ms = (C.c_double * len(OE_KERNELS))()
        nl = (C.c_int32 * len(OE_KERNELS))()
        _check(lib().ukhip_batch_smooth_oddeven_timed(self.h, options, ms, nl), "smooth_timed")
        return dict(zip(OE_KERNELS, list(ms))), dict(zip(OE_KERNELS, list(nl)))

    def smooth_associative_timed(self, options=0):
        """Associative smoothing timed with HIP events after each kernel class
        (ukhip_batch_smooth_associative_timed). Returns ({class: ms},
        {class: launches}) over AS_KERNELS."""
        ms = (C.c_double * len(AS_KERNELS))()
        nl = (C.c_int32 * len(AS_KERNELS))()
        _check(lib().ukhip_batch_smooth_associative_timed(self.h, options, ms, nl), "smooth_associative_timed")
        return dict(zip(AS_KERNELS, list(ms))), dict(zip(AS_KERNELS, list(nl)))

    def timed_parts(self, options=0, mixed=None):
        """ukhip_batch_timed_parts: one eager smoothing (FP64 with `options`, or
        the mixed path with mixed = (refinements, fp32_levels, covariances)),
        parts one after another, HIP events around every launch group. Returns
        [(n, {slot: ms}, {slot: launches})] per part."""
        maxp = 64
        S = len(OE_KERNELS)
        ms = (C.c_double * (maxp * S))()
        nl = (C.c_int32 * (maxp * S))()
        pn = (C.c_int32 * maxp)()
        mo = None if mixed is None else C.byref(MixedOptions(mixed[0], mixed[1], 1 if mixed[2] else 0))
        np_ = lib().ukhip_batch_timed_parts(self.h, options, mo, ms, nl, pn, maxp)
        if np_ < 0:
            raise UkhipError("timed_parts: " + last_error())
        return [(pn[p], dict(zip(OE_KERNELS, ms[p * S:(p + 1) * S])), dict(zip(OE_KERNELS, nl[p * S:(p + 1) * S])))
                for p in range(np_)]

    def download(self, i, want_cov=True):
        p = self.problems[i]
        n = p.n[0]
        x = np.empty((p.k, n))
        cv = np.empty((p.k, n, n)) if want_cov else None
        _check(lib().ukhip_batch_download(self.h, i, x.ctypes.data, cv.ctypes.data if want_cov else None),
               "download")
        if want_cov:
            cv = np.swapaxes(cv, 1, 2)  # column-major per step
        return x, cv

    def download_cov(self, i, fmt):
        """Problem i's covariances as 'C' (k x n x n), 'D' (diagonals, k x n) or
        'W' (inverse Cholesky factors W = L^{-1}, k x n x n, lower triangular)."""
        p = self.problems[i]
        n = p.n[0]
        out = np.empty((p.k, n) if fmt == "D" else (p.k, n, n))
        _check(lib().ukhip_batch_download_cov(self.h, i, fmt.encode(), out.ctypes.data), "download_cov")
        return out if fmt == "D" else np.swapaxes(out, 1, 2)

    def free(self):
        if self.h:
            lib().ukhip_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
