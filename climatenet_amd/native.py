"""ctypes binding of libpdfc.so (the C ABI in include/pdfc.h).

Loading fails loudly when the HIP library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpdfc.so")
_LIB = None

EXPORTS = [
    "pdfc_create", "pdfc_step", "pdfc_get_state", "pdfc_set_state", "pdfc_get_contacts",
    "pdfc_get_self_contacts", "pdfc_get_self_candidates", "pdfc_detect", "pdfc_get_iteration_stats",
    "pdfc_get_phase_stats", "pdfc_get_info", "pdfc_solve", "pdfc_destroy", "pdfc_last_error",
    "pdfc_set_kernel_timing", "pdfc_get_kernel_times",
]


class PdfcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pdfc error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpdfc.so is not built (%s); run python -m climatenet_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        SZ = C.POINTER(C.c_size_t)
        D = C.POINTER(C.c_double)
        L.pdfc_create.argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.Params), C.POINTER(P)]
        L.pdfc_step.argtypes = [P, C.c_int]
        L.pdfc_get_state.argtypes = [P, D, D]
        L.pdfc_set_state.argtypes = [P, D, D]
        L.pdfc_get_contacts.argtypes = [P, C.POINTER(abi.Contact), C.c_size_t, SZ]
        L.pdfc_get_self_contacts.argtypes = [P, C.POINTER(abi.SelfContact), C.c_size_t, SZ]
        L.pdfc_get_self_candidates.argtypes = [P, C.POINTER(abi.Candidate), C.c_size_t, SZ]
        L.pdfc_detect.argtypes = [P]
        L.pdfc_get_iteration_stats.argtypes = [P, D, D, C.c_size_t, SZ]
        L.pdfc_get_phase_stats.argtypes = [P, C.POINTER(abi.PhaseStats)]
        L.pdfc_get_info.argtypes = [P, C.POINTER(abi.Info)]
        L.pdfc_solve.argtypes = [P, D, D]
        L.pdfc_set_kernel_timing.argtypes = [P, C.c_int]
        L.pdfc_get_kernel_times.argtypes = [P, C.POINTER(abi.KernelTimes)]
        L.pdfc_destroy.argtypes = [P]
        L.pdfc_destroy.restype = None
        L.pdfc_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(code):
    if code != abi.PDFC_OK:
        raise PdfcError(code, lib().pdfc_last_error().decode())


class Simulation:
    """One device-resident simulation (SimulationFrictionEstimation on the GPU)."""

    def __init__(self, spec, precision=abi.FP64, refinement_sweeps=2, warm_start=abi.WARM_PLAIN,
                 record_iterations=True, iterations=None):
        self.spec = spec
        self._scene = spec.c_scene()
        self._params = spec.c_params(precision=precision, refinement_sweeps=refinement_sweeps,
                                     warm_start=warm_start, record_iterations=record_iterations,
                                     iterations=iterations)
        h = C.c_void_p()
        _check(lib().pdfc_create(C.byref(self._scene), C.byref(self._params), C.byref(h)))
        self.h = h
        self.n = spec.n_vertices

    def close(self):
        if getattr(self, "h", None):
            lib().pdfc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        _check(lib().pdfc_step(self.h, n))

    def detect(self):
        _check(lib().pdfc_detect(self.h))

    def state(self):
        q = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        _check(lib().pdfc_get_state(self.h, abi.dptr(q), abi.dptr(v)))
        return q, v

    def set_state(self, q=None, v=None):
        q = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().pdfc_set_state(self.h, abi.dptr(q), abi.dptr(v)))

    def _list(self, fn, T):
        n = C.c_size_t()
        code = fn(self.h, None, 0, C.byref(n))
        if code not in (abi.PDFC_OK, abi.PDFC_ERR_CAPACITY):
            _check(code)
        arr = (T * max(1, n.value))()
        _check(fn(self.h, arr, n.value, C.byref(n)))
        return list(arr)[: n.value]

    def contacts(self):
        return self._list(lib().pdfc_get_contacts, abi.Contact)

    def self_contacts(self):
        return self._list(lib().pdfc_get_self_contacts, abi.SelfContact)

    def candidates(self):
        return self._list(lib().pdfc_get_self_candidates, abi.Candidate)

    def iteration_stats(self):
        n = C.c_size_t()
        code = lib().pdfc_get_iteration_stats(self.h, None, None, 0, C.byref(n))
        if code not in (abi.PDFC_OK, abi.PDFC_ERR_CAPACITY):
            _check(code)
        a = np.zeros(max(1, n.value))
        b = np.zeros(max(1, n.value))
        _check(lib().pdfc_get_iteration_stats(self.h, abi.dptr(a), abi.dptr(b), n.value, C.byref(n)))
        return a[: n.value], b[: n.value]

    def info(self):
        I = abi.Info()
        _check(lib().pdfc_get_info(self.h, C.byref(I)))
        out = {}
        for f, _ in abi.Info._fields_:
            v = getattr(I, f)
            out[f] = list(v) if hasattr(v, "__len__") else v
        return out

    def phase_stats(self):
        P = abi.PhaseStats()
        _check(lib().pdfc_get_phase_stats(self.h, C.byref(P)))
        return {f: getattr(P, f) for f, _ in abi.PhaseStats._fields_}

    def set_kernel_timing(self, enable=True):
        _check(lib().pdfc_set_kernel_timing(self.h, int(enable)))

    def kernel_times(self):
        T = abi.KernelTimes()
        _check(lib().pdfc_get_kernel_times(self.h, C.byref(T)))
        return {f: getattr(T, f) for f, _ in abi.KernelTimes._fields_}

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        _check(lib().pdfc_solve(self.h, abi.dptr(b), abi.dptr(x)))
        return x
