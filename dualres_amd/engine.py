"""Python handle over one device chain, mirroring ``MultiResParameters<T>``.

Method names follow include/dualres/MultiResParameters.h (gradient,
log_prior, log_likelihood, log_posterior, update, mu, mu_, sigma,
add_data_to_mu, set_output_indices); arrays on the torus are numpy arrays of
shape (M2, M1, M0) (x fastest, flat index i + M0 (j + M1 k)).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def pow2_embedding(max_ijk):
    """The reference rule M = 2^ceil(log2(2 max_ijk)) (MultiResParameters.h:191-207)."""
    import math
    return tuple(int(2.0 ** math.ceil(math.log(2 * int(m)) / math.log(2.0))) for m in max_ijk)


class Chain:
    """One HMC chain resident on a GPU."""

    def __init__(self, ijk, theta, qform, n_datasets=1, precision="fp64", state_mode="compat",
                 covfun="rbf", embed=None, seed=1, device=0):
        lib = L.load()
        cfg = L.dr_config()
        cfg.n_datasets = int(n_datasets)
        cfg.precision = {"fp64": L.DR_FP64, "fp32": L.DR_FP32}[precision]
        cfg.state_mode = {"compat": L.DR_STATE_COMPAT, "real": L.DR_STATE_REAL}[state_mode]
        cfg.covfun = {"rbf": L.DR_COV_RBF, "rq": L.DR_COV_RQ}[covfun]
        e = (0, 0, 0) if embed is None else tuple(int(v) for v in embed)
        for a in range(3):
            cfg.embed[a] = e[a]
            cfg.theta[a] = float(theta[a])
        q = np.asarray(qform, dtype=np.float32).reshape(16)
        for i in range(16):
            cfg.qform[i] = float(q[i])
        cfg.seed = int(seed) & ((1 << 64) - 1)
        cfg.device = int(device)
        self._ijk = np.ascontiguousarray(ijk, dtype=np.int32)
        assert self._ijk.ndim == 2 and self._ijk.shape[1] == 3
        self.n_h = self._ijk.shape[0]
        self.n_datasets = int(n_datasets)
        self.precision = precision
        self.state_mode = state_mode
        h = C.c_void_p()
        L.check(lib.dr_create(C.byref(cfg), L.iptr32(self._ijk), self.n_h, C.byref(h)))
        self._h = h
        self._lib = lib
        dims = (C.c_int * 3)()
        N = C.c_int64()
        nz = C.c_int64()
        mx = C.c_double()
        L.check(lib.dr_spectrum_report(h, dims, C.byref(N), C.byref(nz), C.byref(mx)))
        self.M = (dims[0], dims[1], dims[2])
        self.N = N.value
        self.n_nonpositive = nz.value
        self.max_eig = mx.value
        self.shape = (self.M[2], self.M[1], self.M[0])
        self.n_out = self.n_h

    def close(self):
        if getattr(self, "_h", None):
            L.check(self._lib.dr_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- data -------------------------------------------------------------
    def set_data(self, Yh, Ys=None, W=None):
        """W: scipy CSR (n_s x n_h) or a (rowptr, cols, vals) triple."""
        Yh = L.as_f64(Yh)
        if self.n_datasets == 2:
            if hasattr(W, "indptr"):
                rp, cols, vals = W.indptr, W.indices, W.data
            else:
                rp, cols, vals = W
            rp = np.ascontiguousarray(rp, dtype=np.int64)
            cols = np.ascontiguousarray(cols, dtype=np.int32)
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            Ys = L.as_f64(Ys)
            L.check(self._lib.dr_set_data(self._h, L.dptr(Yh), Yh.size, L.dptr(Ys), Ys.size, L.iptr64(rp),
                                          L.iptr32(cols), L.dptr(vals)))
        else:
            L.check(self._lib.dr_set_data(self._h, L.dptr(Yh), Yh.size, None, 0, None, None, None))

    def w_layout(self, force_csr):
        """Select plain CSR (True) or the stencil + boundary layout (False, default) before set_data."""
        L.check(self._lib.dr_w_layout(self._h, 1 if force_csr else 0))

    def w_stats(self):
        st = np.zeros(11, dtype=np.int64)
        L.check(self._lib.dr_w_stats(self._h, L.iptr64(st)))
        return {"stencil_layout": bool(st[0]), "interior_rows": int(st[1]), "stencil_entries": int(st[2]),
                "boundary_rows": int(st[3]), "boundary_entries": int(st[4]), "nnz": int(st[5]),
                "touched_voxels": int(st[6]), "interior_tiles": int(st[7]), "tile_cells": int(st[8]),
                "tile_voxels": int(st[9]), "far_rows": int(st[10])}

    def add_data_to_mu(self):
        L.check(self._lib.dr_add_data_to_mu(self._h))

    def set_output_indices(self, ijk):
        ijk = np.ascontiguousarray(ijk, dtype=np.int32)
        L.check(self._lib.dr_set_output_indices(self._h, L.iptr32(ijk), ijk.shape[0]))
        self.n_out = ijk.shape[0]

    def mu_(self):
        out = np.empty(self.n_out)
        L.check(self._lib.dr_get_mu_output(self._h, L.dptr(out), self.n_out))
        return out

    def summary_reset(self):
        L.check(self._lib.dr_summary_reset(self._h))

    def summary_update(self):
        """Add the current state at the output voxels to the device running sums."""
        L.check(self._lib.dr_summary_update(self._h))

    def summary_read(self):
        """(sum mu, sum mu^2, sum |mu| / max|mu|, number of updates) at the output voxels."""
        s1, s2, sa = np.empty(self.n_out), np.empty(self.n_out), np.empty(self.n_out)
        n = C.c_int64()
        L.check(self._lib.dr_summary_read(self._h, L.dptr(s1), L.dptr(s2), L.dptr(sa), self.n_out, C.byref(n)))
        return s1, s2, sa, n.value

    def mu(self):
        out = np.empty(self.n_h)
        L.check(self._lib.dr_get_mu_data(self._h, L.dptr(out), self.n_h))
        return out

    def set_sigma_sq_inv(self, s):
        s = np.ascontiguousarray(np.atleast_1d(s), dtype=np.float64)
        L.check(self._lib.dr_set_sigma_sq_inv(self._h, L.dptr(s)))

    def sigma_sq_inv(self):
        s = np.zeros(2)
        L.check(self._lib.dr_get_sigma_sq_inv(self._h, L.dptr(s)))
        return s[: self.n_datasets]

    def set_tau_sq_inv(self, t):
        """_update_tau_sq_inv's assignment: lambda *= t / (previous tau^-2)."""
        L.check(self._lib.dr_set_tau_sq_inv(self._h, float(t)))

    def tau_sq_inv(self):
        t = C.c_double()
        L.check(self._lib.dr_get_tau_sq_inv(self._h, C.byref(t)))
        return t.value

    def mu_data_sumsq(self):
        """Sum of Re(mu)^2 over the data voxels of the current state."""
        s = C.c_double()
        L.check(self._lib.dr_mu_data_sumsq(self._h, C.byref(s)))
        return s.value

    def lambda_(self):
        out = np.empty(self.N)
        L.check(self._lib.dr_get_lambda(self._h, L.dptr(out)))
        return out.reshape(self.shape)

    # -- energies ---------------------------------------------------------
    def _scalar(self, fn):
        v = C.c_double()
        L.check(fn(self._h, C.byref(v)))
        return v.value

    def log_prior(self):
        return self._scalar(self._lib.dr_log_prior)

    def log_likelihood(self):
        return self._scalar(self._lib.dr_log_likelihood)

    def log_posterior(self):
        return self._scalar(self._lib.dr_log_posterior)

    def sse(self):
        a, b = C.c_double(), C.c_double()
        L.check(self._lib.dr_sse(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- HMC --------------------------------------------------------------
    def update(self, eps, L_steps, u_accept, noise=None):
        res = L.dr_update_result()
        if noise is None:
            L.check(self._lib.dr_update(self._h, float(eps), int(L_steps), float(u_accept), C.byref(res)))
        else:
            zr = L.as_f64(np.asarray(noise[0]).reshape(-1))
            zi = L.as_f64(np.asarray(noise[1]).reshape(-1)) if noise[1] is not None else None
            L.check(self._lib.dr_update_with_noise(self._h, float(eps), int(L_steps), float(u_accept),
                                                   L.dptr(zr), L.dptr(zi), C.byref(res)))
        return res.as_dict()

    def autotune(self, cache=None):
        """Pick the FFT pass shapes by timing (or from the JSON cache file);
        returns (params, from_cache)."""
        prm = (C.c_int * 8)()
        hit = C.c_int()
        path = None if cache is None else str(cache).encode()
        L.check(self._lib.dr_autotune(self._h, path, prm, C.byref(hit)))
        return list(prm), bool(hit.value)

    def set_tuning(self, params):
        prm = (C.c_int * 8)(*[int(v) for v in params])
        L.check(self._lib.dr_set_tuning(self._h, prm))

    def tuning(self):
        prm = (C.c_int * 8)()
        L.check(self._lib.dr_get_tuning(self._h, prm))
        return list(prm)

    def update_captured(self, eps, L_steps, u_accept):
        """update() through the chain's captured HIP graph for this leapfrog length."""
        res = L.dr_update_result()
        L.check(self._lib.dr_update_captured(self._h, float(eps), int(L_steps), float(u_accept), C.byref(res)))
        return res.as_dict()

    def leapfrog(self, eps, L_steps):
        res = L.dr_update_result()
        L.check(self._lib.dr_leapfrog(self._h, float(eps), int(L_steps), C.byref(res)))
        return res.as_dict()

    # -- state hooks --------------------------------------------------------
    def set_state(self, mu, p):
        """mu, p: complex arrays on the torus (imaginary parts used in compat mode)."""
        mu = np.asarray(mu).reshape(-1)
        p = np.asarray(p).reshape(-1)
        mr, mi = L.as_f64(mu.real), L.as_f64(np.imag(mu))
        pr, pi = L.as_f64(p.real), L.as_f64(np.imag(p))
        L.check(self._lib.dr_set_state(self._h, L.dptr(mr), L.dptr(mi), L.dptr(pr), L.dptr(pi)))

    def get_state(self):
        mr, mi, pr, pi = (np.empty(self.N) for _ in range(4))
        L.check(self._lib.dr_get_state(self._h, L.dptr(mr), L.dptr(mi), L.dptr(pr), L.dptr(pi)))
        return (mr + 1j * mi).reshape(self.shape), (pr + 1j * pi).reshape(self.shape)

    def gradient(self):
        gr, gi = np.empty(self.N), np.empty(self.N)
        L.check(self._lib.dr_gradient(self._h, L.dptr(gr), L.dptr(gi)))
        return (gr + 1j * gi).reshape(self.shape)

    def profile(self, enable=True):
        L.check(self._lib.dr_profile_enable(self._h, 1 if enable else 0))

    def profile_kernels(self):
        """Per-kernel ms of one leapfrog step (launch order) and the step count."""
        ms = np.zeros(16)
        nk, ns = C.c_int(), C.c_int64()
        L.check(self._lib.dr_profile_kernels(self._h, L.dptr(ms), 16, C.byref(nk), C.byref(ns)))
        names = ["z_fft_kick_drift", "c2r_y", "c2r_x"]
        if self.n_datasets == 2:
            names += ["W_mu", "Wt_v"]
        names += ["r2c_x_residual", "r2c_y"]
        return dict(zip(names, ms[: nk.value].tolist())), ns.value

    def profile_read(self):
        ms, n, a, b = C.c_double(), C.c_int64(), C.c_double(), C.c_double()
        L.check(self._lib.dr_profile_read(self._h, C.byref(ms), C.byref(n), C.byref(a), C.byref(b)))
        return ms.value, n.value
