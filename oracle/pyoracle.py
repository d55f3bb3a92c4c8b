"""ctypes loader of the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, never by the climatenet_amd package.
"""
import ctypes as C
import os

import numpy as np

from climatenet_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.Params), C.POINTER(P)]
        L.orc_destroy.argtypes = [P]
        L.orc_destroy.restype = None
        L.orc_step.argtypes = [P, C.c_int]
        L.orc_set_warm.argtypes = [P, C.POINTER(C.c_double), C.c_size_t, C.c_int]
        L.orc_set_warm.restype = None
        L.orc_set_iterations.argtypes = [P, C.c_int]
        L.orc_set_iterations.restype = None
        L.orc_detect.argtypes = [P, C.c_int]
        L.orc_get_state.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_set_state.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_get_contacts.argtypes = [P, C.POINTER(abi.Contact), C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_get_self_contacts.argtypes = [P, C.POINTER(abi.SelfContact), C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_get_self_candidates.argtypes = [P, C.POINTER(abi.Candidate), C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_get_iteration_stats.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_size_t)]
        L.orc_solve.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_get_sizes.argtypes = [P, C.POINTER(C.c_int64)]
        L.orc_get_K.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.orc_get_masses.argtypes = [P, C.POINTER(C.c_double)]
        L.orc_get_springs.argtypes = [P, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_get_bending_row.argtypes = [P, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int]
        L.orc_factor_seconds.argtypes = [P]
        L.orc_factor_seconds.restype = C.c_double
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_set_num_threads.restype = None
        L.orc_num_threads.restype = C.c_int
        L.orc_last_error.restype = C.c_char_p
        L.orc_mean_value_coefficients.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_int, C.POINTER(C.c_double)]
        L.orc_mean_value_coefficients.restype = None
        L.orc_closest_point.argtypes = [C.POINTER(C.c_double)] * 4 + [C.c_double] + [C.POINTER(C.c_double)] * 2
        L.orc_closest_point.restype = None
        L.orc_cone.argtypes = [C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
        L.orc_alart_curnier.argtypes = [C.POINTER(C.c_double)] * 2 + [C.c_double, C.POINTER(C.c_double)]
        L.orc_alart_curnier.restype = C.c_double
        L.orc_contact_basis.argtypes = [C.POINTER(C.c_double)] * 2
        L.orc_contact_basis.restype = None
        L.orc_pair_candidates.argtypes = [P, C.POINTER(abi.Candidate), C.c_size_t]
        I64 = C.POINTER(C.c_int64)
        I32 = C.POINTER(C.c_int32)
        L.orc_contact_levels.argtypes = [C.c_int, I64, I64, C.c_size_t, I64, C.c_size_t, I32, I32]
        F = C.POINTER(C.c_float)
        L.orc_mlp_forward.argtypes = [F, F, F, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.orc_mlp_forward.restype = None
        L.orc_mlp_gradient.argtypes = [F, F, F, F, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.orc_mlp_gradient.restype = C.c_double
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """The CPU restatement of the reference simulation on one scene."""

    def __init__(self, spec, iterations=None):
        self.spec = spec
        L = lib()
        self._scene = spec.c_scene()
        self._params = spec.c_params(iterations=iterations)
        h = C.c_void_p()
        if L.orc_create(C.byref(self._scene), C.byref(self._params), C.byref(h)) != 0:
            raise ValueError(L.orc_last_error().decode())
        self.h = h
        self.n = spec.n_vertices

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        if lib().orc_step(self.h, n) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def detect(self, brute=False):
        if lib().orc_detect(self.h, int(brute)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def set_warm(self, r_local, hold=1):
        """Warm start of the next step: reactions [C x 3] in that step's contact order."""
        r = np.ascontiguousarray(r_local, dtype=np.float64).reshape(-1, 3)
        lib().orc_set_warm(self.h, _d(r), r.shape[0], int(hold))

    def set_iterations(self, K):
        lib().orc_set_iterations(self.h, int(K))

    def state(self):
        q = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        lib().orc_get_state(self.h, _d(q), _d(v))
        return q, v

    def set_state(self, q=None, v=None):
        lib().orc_set_state(self.h, _d(q) if q is not None else None, _d(v) if v is not None else None)

    def _list(self, fn, T):
        n = C.c_size_t()
        fn(self.h, None, 0, C.byref(n))
        arr = (T * max(1, n.value))()
        fn(self.h, arr, n.value, C.byref(n))
        return list(arr)[: n.value]

    def contacts(self):
        return self._list(lib().orc_get_contacts, abi.Contact)

    def self_contacts(self):
        return self._list(lib().orc_get_self_contacts, abi.SelfContact)

    def candidates(self):
        return self._list(lib().orc_get_self_candidates, abi.Candidate)

    def iteration_stats(self):
        n = C.c_size_t()
        lib().orc_get_iteration_stats(self.h, None, None, 0, C.byref(n))
        a = np.zeros(max(1, n.value))
        b = np.zeros(max(1, n.value))
        lib().orc_get_iteration_stats(self.h, _d(a), _d(b), n.value, C.byref(n))
        return a[: n.value], b[: n.value]

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().orc_solve(self.h, _d(b), _d(x))
        return x

    def sizes(self):
        s = (C.c_int64 * 6)()
        lib().orc_get_sizes(self.h, s)
        return dict(zip(["n", "T", "E", "B", "nnz_K", "nnz_L"], list(s)))

    def K(self):
        import scipy.sparse as sp
        sz = self.sizes()
        ptr = np.zeros(self.n + 1, dtype=np.int64)
        idx = np.zeros(sz["nnz_K"], dtype=np.int32)
        val = np.zeros(sz["nnz_K"])
        lib().orc_get_K(self.h, ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                        idx.ctypes.data_as(C.POINTER(C.c_int32)), _d(val))
        return sp.csr_matrix((val, idx, ptr), shape=(self.n, self.n))

    def masses(self):
        m = np.zeros(self.n)
        lib().orc_get_masses(self.h, _d(m))
        return m

    def springs(self):
        E = self.sizes()["E"]
        ij = np.zeros(2 * E, dtype=np.int32)
        w = np.zeros(E)
        L0 = np.zeros(E)
        lib().orc_get_springs(self.h, ij.ctypes.data_as(C.POINTER(C.c_int32)), _d(w), _d(L0))
        return ij.reshape(-1, 2), w, L0

    def bending_row(self, b):
        idx = np.zeros(32, dtype=np.int32)
        cf = np.zeros(32)
        c = lib().orc_get_bending_row(self.h, b, idx.ctypes.data_as(C.POINTER(C.c_int32)), _d(cf), 32)
        return idx[:c], cf[:c]

    def factor_seconds(self):
        return lib().orc_factor_seconds(self.h)

    def pair_candidates(self, cands):
        """Greedy pairing + level ordering of given candidates [(vertex, triangle, (b0, b1, b2))]
        at the current positions; returns the self contacts."""
        arr = (abi.Candidate * max(len(cands), 1))()
        for i, (v, t, b) in enumerate(cands):
            arr[i].vertex = v
            arr[i].triangle = t
            for k in range(3):
                arr[i].barycentric[k] = b[k]
        if lib().orc_pair_candidates(self.h, arr, len(cands)) != 0:
            raise ValueError(lib().orc_last_error().decode())
        return self.self_contacts()


def mean_value_coefficients(center, ring):
    center = np.ascontiguousarray(center, dtype=np.float64)
    ring = np.ascontiguousarray(ring, dtype=np.float64).ravel()
    out = np.zeros(ring.size // 3)
    lib().orc_mean_value_coefficients(_d(center), _d(ring), out.size, _d(out))
    return out


def closest_point(p, a, b, c, tol2):
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (p, a, b, c)]
    cp = np.zeros(3)
    bary = np.zeros(3)
    lib().orc_closest_point(*[_d(x) for x in args], float(tol2), _d(cp), _d(bary))
    return cp, bary


def cone(f, mu):
    f = np.ascontiguousarray(f, dtype=np.float64)
    r = np.zeros(3)
    act = lib().orc_cone(_d(f), float(mu), _d(r))
    return bool(act), r


def alart_curnier(u, r, mu):
    u = np.ascontiguousarray(u, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    ac = np.zeros(3)
    nrm = lib().orc_alart_curnier(_d(u), _d(r), float(mu), _d(ac))
    return nrm, ac


def contact_levels(n_vertices, pairs, seeds=()):
    """Level ordering of self-contact pairs [(vertex, partner)] seeded by obstacle-contact
    vertices: (levels as lists of pair indices in computation order, level of each pair)."""
    pr = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    v = np.ascontiguousarray(pr[:, 0])
    w = np.ascontiguousarray(pr[:, 1])
    sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64).reshape(-1))
    lev = np.full(len(pr), -1, dtype=np.int32)
    order = np.full(len(pr), -1, dtype=np.int32)
    I64 = C.POINTER(C.c_int64)
    I32 = C.POINTER(C.c_int32)
    nl = lib().orc_contact_levels(int(n_vertices), v.ctypes.data_as(I64), w.ctypes.data_as(I64), len(pr),
                                  sd.ctypes.data_as(I64), sd.size, lev.ctypes.data_as(I32), order.ctypes.data_as(I32))
    if nl < 0:
        raise ValueError(lib().orc_last_error().decode())
    levels = [[] for _ in range(nl)]
    for s in order:
        levels[lev[s]].append(int(s))
    return levels, lev


def contact_basis(n):
    n = np.ascontiguousarray(n, dtype=np.float64)
    B = np.zeros(9)
    lib().orc_contact_basis(_d(n), _d(B))
    return B.reshape(3, 3)  # rows: n, t, b


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_forward(params, norm, x, bf16=True):
    """FP64 contact MLP (with the device's BF16 rounding points when bf16): y [rows x 3]."""
    P = np.ascontiguousarray(params, dtype=np.float32)
    Z = np.ascontiguousarray(norm, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
    y = np.zeros((x.shape[0], 3))
    lib().orc_mlp_forward(_f(P), _f(Z), _f(x), x.shape[0], int(bf16), _d(y))
    return y


def mlp_gradient(params, norm, x, t, bf16=True):
    """(gradient [10448], loss) of the mean squared error in normalised output units."""
    P = np.ascontiguousarray(params, dtype=np.float32)
    Z = np.ascontiguousarray(norm, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 12)
    t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 3)
    g = np.zeros(10448)
    loss = lib().orc_mlp_gradient(_f(P), _f(Z), _f(x), _f(t), x.shape[0], int(bf16), _d(g))
    return g, loss


def num_threads():
    return int(lib().orc_num_threads())


def set_num_threads(k):
    """OpenMP threads of the oracle's parallel regions (all threads by default)."""
    lib().orc_set_num_threads(int(k))
