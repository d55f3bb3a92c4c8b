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
    "pdfc_create", "pdfc_create_batch", "pdfc_step", "pdfc_get_state", "pdfc_set_state", "pdfc_get_contacts",
    "pdfc_get_self_contacts", "pdfc_get_self_candidates", "pdfc_detect", "pdfc_get_iteration_stats",
    "pdfc_get_phase_stats", "pdfc_get_info", "pdfc_solve", "pdfc_destroy", "pdfc_last_error",
    "pdfc_set_kernel_timing", "pdfc_get_kernel_times", "pdfc_set_mlp", "pdfc_get_warm_data",
    "pdfc_set_warm_reactions", "pdfc_save_checkpoint", "pdfc_load_checkpoint",
    "pdfc_mlp_create", "pdfc_mlp_destroy", "pdfc_mlp_set_params", "pdfc_mlp_get_params",
    "pdfc_mlp_forward", "pdfc_mlp_gradient", "pdfc_mlp_train", "pdfc_mlp_gradient_noalloc",
    "pdfc_step_begin", "pdfc_step_iterate", "pdfc_step_end",
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
        L.pdfc_create_batch.argtypes = [C.POINTER(abi.Scene), C.c_size_t, C.POINTER(abi.Params), C.POINTER(P)]
        L.pdfc_step.argtypes = [P, C.c_int]
        L.pdfc_step_begin.argtypes = [P]
        L.pdfc_step_iterate.argtypes = [P, C.c_int]
        L.pdfc_step_end.argtypes = [P]
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
        F = C.POINTER(C.c_float)
        L.pdfc_set_mlp.argtypes = [P, P]
        L.pdfc_set_warm_reactions.argtypes = [P, C.POINTER(abi.Contact), C.c_size_t, C.c_int]
        L.pdfc_get_warm_data.argtypes = [P, F, F, C.c_size_t, SZ]
        L.pdfc_save_checkpoint.argtypes = [P, C.c_void_p, C.c_size_t, SZ]
        L.pdfc_load_checkpoint.argtypes = [P, C.c_void_p, C.c_size_t]
        L.pdfc_mlp_create.argtypes = [C.POINTER(P), C.c_uint64]
        L.pdfc_mlp_destroy.argtypes = [P]
        L.pdfc_mlp_destroy.restype = None
        L.pdfc_mlp_set_params.argtypes = [P, F, F]
        L.pdfc_mlp_get_params.argtypes = [P, F, F]
        L.pdfc_mlp_forward.argtypes = [P, F, C.c_size_t, F]
        L.pdfc_mlp_gradient.argtypes = [P, F, F, C.c_size_t, F, F]
        L.pdfc_mlp_train.argtypes = [P, F, F, C.c_size_t, C.c_int, C.c_int, C.c_float, C.c_uint64, F]
        L.pdfc_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(code):
    if code != abi.PDFC_OK:
        raise PdfcError(code, lib().pdfc_last_error().decode())


class Simulation:
    """One device-resident simulation (SimulationFrictionEstimation on the GPU)."""

    def __init__(self, spec, precision=abi.FP64, refinement_sweeps=2, warm_start=abi.WARM_PLAIN,
                 record_iterations=True, iterations=None, residual_tolerance=0.0, max_iterations=1000,
                 mlp=None, use_graph=False, warm_hold=1):
        self.spec = spec
        self._scene = spec.c_scene()
        self._params = spec.c_params(precision=precision, refinement_sweeps=refinement_sweeps,
                                     warm_start=warm_start, record_iterations=record_iterations,
                                     iterations=iterations, residual_tolerance=residual_tolerance,
                                     max_iterations=max_iterations, use_graph=use_graph, warm_hold=warm_hold)
        h = C.c_void_p()
        _check(lib().pdfc_create(C.byref(self._scene), C.byref(self._params), C.byref(h)))
        self.h = h
        self.n = spec.n_vertices
        self._mlp = None
        if mlp is not None:
            self.set_mlp(mlp)

    def set_mlp(self, mlp):
        """Attach a ContactMLP (kept alive by this object) for the learned warm start."""
        _check(lib().pdfc_set_mlp(self.h, None if mlp is None else mlp.h))
        self._mlp = mlp

    def set_warm_reactions(self, contacts, hold=1):
        """Warm start of the next step from pdfc_contact records (e.g. a converged step's)."""
        arr = (abi.Contact * max(1, len(contacts)))(*contacts)
        _check(lib().pdfc_set_warm_reactions(self.h, arr, len(contacts), int(hold)))

    def warm_data(self):
        """(features [C x 12], targets [C x 3]) of the last step's obstacle contacts."""
        n = C.c_size_t()
        code = lib().pdfc_get_warm_data(self.h, None, None, 0, C.byref(n))
        if code not in (abi.PDFC_OK, abi.PDFC_ERR_CAPACITY):
            _check(code)
        x = np.zeros((max(1, n.value), abi.MLP_FEATURES), dtype=np.float32)
        y = np.zeros((max(1, n.value), abi.MLP_OUTPUTS), dtype=np.float32)
        _check(lib().pdfc_get_warm_data(self.h, abi.fptr(x), abi.fptr(y), n.value, C.byref(n)))
        return x[: n.value], y[: n.value]

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

    def step_phases(self, iterations=None):
        """One step through pdfc_step_begin / pdfc_step_iterate / pdfc_step_end (the
        reference's hook sequence), `iterations` iterations (default: the scene's)."""
        k = self.spec.iterations if iterations is None else int(iterations)
        _check(lib().pdfc_step_begin(self.h))
        for _ in range(k):
            _check(lib().pdfc_step_iterate(self.h, 1))
        _check(lib().pdfc_step_end(self.h))

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

    def save_checkpoint(self) -> bytes:
        n = C.c_size_t()
        code = lib().pdfc_save_checkpoint(self.h, None, 0, C.byref(n))
        if code != abi.PDFC_ERR_CAPACITY:
            _check(code)
        buf = C.create_string_buffer(n.value)
        _check(lib().pdfc_save_checkpoint(self.h, buf, n.value, C.byref(n)))
        return buf.raw[: n.value]

    def load_checkpoint(self, data: bytes):
        buf = C.create_string_buffer(bytes(data), len(data))
        _check(lib().pdfc_load_checkpoint(self.h, buf, len(data)))

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

    def tuning(self) -> str:
        """The chosen substitution schedule as a PDFC_TUNE value: another simulation of the same
        scene created with it set sums in the same order (bit-identical results)."""
        I = self.info()
        nl = I["n_levels_swept"]
        if nl > len(I["waves_forward"]):
            raise ValueError("more levels than pdfc_info reports")
        f = ",".join(str(w) for w in I["waves_forward"][:nl])
        b = ",".join(str(w) for w in I["waves_backward"][:nl])
        g = ",".join(str(x) for x in I["parts_forward"][:I["launches_forward"]])
        h = ",".join(str(x) for x in I["parts_backward"][:I["launches_backward"]])
        # t=0: symmetric tiles (or no top block); o: the top block's finish folded or not;
        # g, h: levels per launch of each sweep (merged ranges)
        out = "f=%s;b=%s;t=%d;o=%d;q=%d" % (f, b, I["top_chunks"], I["top_fold"], I["top_pregather"])
        if nl > 0:
            out += ";g=%s;h=%s;p=%s" % (g, h, ",".join(str(x) for x in I["split_backward"][:nl]))
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


class BatchSimulation(Simulation):
    """Independent scenes advanced together (one block-diagonal system, one kernel sequence
    per step). Per-vertex arrays concatenate the scenes in order; ``ranges[k]`` is scene k's
    vertex range."""

    def __init__(self, specs, precision=abi.FP64, refinement_sweeps=2, warm_start=abi.WARM_PLAIN,
                 record_iterations=True, iterations=None, residual_tolerance=0.0, max_iterations=1000,
                 use_graph=False):
        if not specs:
            raise ValueError("no scene")
        for sp in specs[1:]:
            if sp.time_step != specs[0].time_step or sp.iterations != specs[0].iterations:
                raise ValueError("batched scenes share the time step and iteration count")
        self.specs = list(specs)
        self.spec = specs[0]
        self._scenes = (abi.Scene * len(specs))()
        for k, sp in enumerate(specs):
            self._scenes[k] = sp.c_scene()
        self._params = specs[0].c_params(precision=precision, refinement_sweeps=refinement_sweeps,
                                         warm_start=warm_start, record_iterations=record_iterations,
                                         iterations=iterations, residual_tolerance=residual_tolerance,
                                         max_iterations=max_iterations, use_graph=use_graph)
        h = C.c_void_p()
        _check(lib().pdfc_create_batch(self._scenes, len(specs), C.byref(self._params), C.byref(h)))
        self.h = h
        off = np.cumsum([0] + [sp.n_vertices for sp in specs])
        self.ranges = [(int(off[k]), int(off[k + 1])) for k in range(len(specs))]
        self.n = int(off[-1])
        self._mlp = None


class ContactMLP:
    """The contact network of the learned warm start (12 -> 64 -> 64 -> 64 -> 3, SiLU, BF16 MFMA)."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        _check(lib().pdfc_mlp_create(C.byref(h), seed))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pdfc_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        z = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        if p is not None and p.size != abi.MLP_PARAMS:
            raise ValueError("expected %d parameters" % abi.MLP_PARAMS)
        if z is not None and z.size != abi.MLP_NORM:
            raise ValueError("expected %d normalisation values" % abi.MLP_NORM)
        _check(lib().pdfc_mlp_set_params(self.h, abi.fptr(p), abi.fptr(z)))

    def params(self):
        p = np.zeros(abi.MLP_PARAMS, dtype=np.float32)
        z = np.zeros(abi.MLP_NORM, dtype=np.float32)
        _check(lib().pdfc_mlp_get_params(self.h, abi.fptr(p), abi.fptr(z)))
        return p, z

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, abi.MLP_FEATURES)
        y = np.zeros((x.shape[0], abi.MLP_OUTPUTS), dtype=np.float32)
        _check(lib().pdfc_mlp_forward(self.h, abi.fptr(x), x.shape[0], abi.fptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, abi.MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, abi.MLP_OUTPUTS)
        g = np.zeros(abi.MLP_PARAMS, dtype=np.float32)
        loss = C.c_float()
        _check(lib().pdfc_mlp_gradient(self.h, abi.fptr(x), abi.fptr(t), x.shape[0], abi.fptr(g),
                                       C.cast(C.byref(loss), C.POINTER(C.c_float))))
        return g, float(loss.value)

    def train(self, x, t, epochs, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, abi.MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, abi.MLP_OUTPUTS)
        losses = np.zeros(epochs, dtype=np.float32)
        _check(lib().pdfc_mlp_train(self.h, abi.fptr(x), abi.fptr(t), x.shape[0], epochs, batch, lr, seed,
                                    abi.fptr(losses)))
        return losses
