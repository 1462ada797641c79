"""ctypes handle on the C++ CPU restatement of the reference's sampler
(oracle/cpu_hmc.cpp, built as oracle/libdr_cpu_port.so; test and CPU-baseline
infrastructure only).  Arrays on the torus are numpy arrays of shape
(M2, M1, M0), complex, in real space as the reference keeps its state."""
import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdr_cpu_port.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is not built (run make -C oracle)")
        lib = C.CDLL(LIB)
        P, D, I32, I64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        lib.drc_create.argtypes = [C.c_int, C.c_int, D, C.POINTER(C.c_float), C.POINTER(C.c_int), I32, C.c_int64, D,
                                   D, C.c_int64, I64, I32, D, C.c_int, C.c_int, C.c_uint32, C.POINTER(P)]
        for name, args in (("drc_destroy", [P]), ("drc_nonpositive", [P, I64]), ("drc_set_sigma", [P, D]),
                           ("drc_set_state", [P, D, D, D, D]), ("drc_get_state", [P, D, D, D, D]),
                           ("drc_gradient", [P, D, D]), ("drc_leapfrog", [P, C.c_double, C.c_int, D]),
                           ("drc_update", [P, C.c_double, C.c_int, C.c_double, D, D, D]),
                           ("drc_add_data_to_mu", [P])):
            getattr(lib, name).argtypes = args
        for name in ("drc_create", "drc_destroy", "drc_nonpositive", "drc_set_sigma", "drc_set_state",
                     "drc_get_state", "drc_gradient", "drc_leapfrog", "drc_update", "drc_add_data_to_mu"):
            getattr(lib, name).restype = C.c_int
        lib.drc_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class CpuChain:
    """MultiResParameters<T> on the host cores (T = float64 or float32)."""

    def __init__(self, ijk, theta, qform, M, Yh, Ys=None, W=None, precision="fp64", threads=1, w_threads=1,
                 seed=42):
        lib = load()
        self.M = tuple(int(m) for m in M)
        self.N = int(np.prod(self.M))
        self.shape = (self.M[2], self.M[1], self.M[0])
        ijk = np.ascontiguousarray(ijk, dtype=np.int32)
        Yh = np.ascontiguousarray(Yh, dtype=np.float64)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        q = np.ascontiguousarray(qform, dtype=np.float32).reshape(16)
        m = (C.c_int * 3)(*self.M)
        nd = 1
        rp = cols = vals = ys = None
        n_s = 0
        if W is not None:
            nd = 2
            rp = np.ascontiguousarray(W.indptr, dtype=np.int64)
            cols = np.ascontiguousarray(W.indices, dtype=np.int32)
            vals = np.ascontiguousarray(W.data, dtype=np.float64)
            ys = np.ascontiguousarray(Ys, dtype=np.float64)
            n_s = ys.size
        self._keep = (ijk, Yh, th, q, rp, cols, vals, ys)
        h = C.c_void_p()
        rc = lib.drc_create(0 if precision == "fp64" else 1, nd, _d(th), q.ctypes.data_as(C.POINTER(C.c_float)), m,
                            ijk.ctypes.data_as(C.POINTER(C.c_int32)), ijk.shape[0], _d(Yh), _d(ys), n_s,
                            None if rp is None else rp.ctypes.data_as(C.POINTER(C.c_int64)),
                            None if cols is None else cols.ctypes.data_as(C.POINTER(C.c_int32)), _d(vals),
                            int(threads), int(w_threads), int(seed) & 0xFFFFFFFF, C.byref(h))
        self._check(rc)
        self._h = h
        self.nd = nd

    def _check(self, rc):
        if rc != 0:
            msg = load().drc_last_error().decode()
            raise (ValueError if rc == 1 else RuntimeError)(msg)

    def close(self):
        if getattr(self, "_h", None):
            load().drc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def nonpositive(self):
        n = C.c_int64()
        load().drc_nonpositive(self._h, C.byref(n))
        return n.value

    def set_sigma_sq_inv(self, s):
        s = np.ascontiguousarray(np.atleast_1d(s), dtype=np.float64)
        self._check(load().drc_set_sigma(self._h, _d(s)))

    def set_state(self, mu, p):
        mu = np.asarray(mu).reshape(-1)
        p = np.asarray(p).reshape(-1)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (mu.real, np.imag(mu), p.real, np.imag(p))]
        self._check(load().drc_set_state(self._h, *[_d(x) for x in a]))

    def get_state(self):
        a = [np.empty(self.N) for _ in range(4)]
        self._check(load().drc_get_state(self._h, *[_d(x) for x in a]))
        return (a[0] + 1j * a[1]).reshape(self.shape), (a[2] + 1j * a[3]).reshape(self.shape)

    def gradient(self):
        gr, gi = np.empty(self.N), np.empty(self.N)
        self._check(load().drc_gradient(self._h, _d(gr), _d(gi)))
        return (gr + 1j * gi).reshape(self.shape)

    def leapfrog(self, eps, L):
        """(log posterior, kinetic) before and after L steps; the end point becomes the state."""
        out = np.zeros(4)
        self._check(load().drc_leapfrog(self._h, float(eps), int(L), _d(out)))
        return out

    def update(self, eps, L, u=-1.0, noise=None):
        """One HMC iteration with the Gibbs draws; returns (H0, H1, R, accepted)."""
        out = np.zeros(4)
        zr = zi = None
        if noise is not None:
            zr = np.ascontiguousarray(np.asarray(noise[0]).reshape(-1), dtype=np.float64)
            zi = np.ascontiguousarray(np.asarray(noise[1]).reshape(-1), dtype=np.float64)
        self._check(load().drc_update(self._h, float(eps), int(L), float(u), _d(zr), _d(zi), _d(out)))
        return out

    def add_data_to_mu(self):
        self._check(load().drc_add_data_to_mu(self._h))
