"""Python view of the fit server in libdualres_host.so (dualres_amd/host/server.hpp).

A Server holds worker threads that run whole posterior fits (the fit_model
loop of include/dualres/gaussian_process_model.h:87-177, with the kriging
build for dual-resolution requests) on one device, admitting queued
requests while their device bytes fit the budget.  Volumes are numpy arrays
of shape (nz, ny, nx) (x fastest), qforms row-major 4x4.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class _Request(C.Structure):
    _fields_ = [
        ("n_datasets", C.c_int),
        ("hr", C.POINTER(C.c_double)),
        ("hr_dims", C.c_int * 3),
        ("q_hr", C.c_float * 16),
        ("sr", C.POINTER(C.c_double)),
        ("sr_dims", C.c_int * 3),
        ("q_sr", C.c_float * 16),
        ("theta", C.c_double * 3),
        ("neighborhood", C.c_double),
        ("burnin", C.c_int), ("nsave", C.c_int), ("thin", C.c_int), ("L", C.c_int),
        ("mh_target", C.c_double),
        ("seed", C.c_uint64),
        ("precision", C.c_int), ("state_mode", C.c_int),
        ("embed", C.c_int * 3),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("mh_rate", C.c_double), ("eps", C.c_double),
        ("queue_s", C.c_double), ("run_s", C.c_double), ("total_s", C.c_double),
        ("grad_evals", C.c_int64), ("device_bytes", C.c_int64), ("n_out", C.c_int64),
    ]


def load_host():
    path = os.path.join(_HERE, "libdualres_host.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (make -C dualres_amd/host)")
    lib = C.CDLL(path)
    lib.drh_server_create.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    lib.drh_server_submit.argtypes = [C.c_void_p, C.POINTER(_Request), C.POINTER(C.c_int64)]
    lib.drh_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_double, C.POINTER(_Result),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64, C.c_char_p, C.c_int]
    lib.drh_server_destroy.argtypes = [C.c_void_p]
    lib.drh_server_cancel.argtypes = [C.c_void_p, C.c_int64]
    lib.drh_server_cancel.restype = C.c_int
    lib.drh_estimate_request_bytes.argtypes = [C.POINTER(_Request), C.POINTER(C.c_int64)]
    return lib


def make_request(hr, q_hr, theta, std=None, q_std=None, neighborhood=-1.0, burnin=100, nsave=100, thin=1, L=25,
                 mh_target=0.65, seed=1, precision="fp32", state_mode="compat", embed=(0, 0, 0)):
    """Request dict; hr/std are (nz, ny, nx) volumes (0 or NaN = missing)."""
    return {"hr": np.ascontiguousarray(hr, dtype=np.float64), "q_hr": np.asarray(q_hr, np.float32).reshape(16),
            "std": None if std is None else np.ascontiguousarray(std, dtype=np.float64),
            "q_std": None if q_std is None else np.asarray(q_std, np.float32).reshape(16),
            "theta": tuple(float(t) for t in theta), "neighborhood": float(neighborhood), "burnin": int(burnin),
            "nsave": int(nsave), "thin": int(thin), "L": int(L), "mh_target": float(mh_target), "seed": int(seed),
            "precision": precision, "state_mode": state_mode, "embed": tuple(int(e) for e in embed)}


def _to_struct(r):
    q = _Request()
    q.n_datasets = 2 if r["std"] is not None else 1
    q.hr = r["hr"].ctypes.data_as(C.POINTER(C.c_double))
    nz, ny, nx = r["hr"].shape
    q.hr_dims[:] = [nx, ny, nz]
    q.q_hr[:] = [float(v) for v in r["q_hr"]]
    if r["std"] is not None:
        q.sr = r["std"].ctypes.data_as(C.POINTER(C.c_double))
        sz, sy, sx = r["std"].shape
        q.sr_dims[:] = [sx, sy, sz]
        q.q_sr[:] = [float(v) for v in r["q_std"]]
    q.theta[:] = list(r["theta"])
    q.neighborhood = r["neighborhood"]
    q.burnin, q.nsave, q.thin, q.L = r["burnin"], r["nsave"], r["thin"], r["L"]
    q.mh_target = r["mh_target"]
    q.seed = r["seed"]
    q.precision = 0 if r["precision"] == "fp64" else 1
    q.state_mode = 0 if r["state_mode"] == "compat" else 1
    q.embed[:] = list(r["embed"])
    return q


def n_outputs(r):
    v = r["hr"]
    return int(np.count_nonzero(~(np.isnan(v) | (v == 0))))


def estimate_bytes(r):
    lib = load_host()
    b = C.c_int64()
    if lib.drh_estimate_request_bytes(C.byref(_to_struct(r)), C.byref(b)) != 0:
        raise RuntimeError("drh_estimate_request_bytes failed")
    return b.value


class Server:
    """Fit server: ``submit`` returns an id, ``wait`` the result dict."""

    def __init__(self, budget_bytes, workers=4, device=0, backfill=True):
        self._lib = load_host()
        self._h = C.c_void_p()
        if self._lib.drh_server_create(int(budget_bytes), int(workers), int(device), int(bool(backfill)),
                                       C.byref(self._h)) != 0:
            raise RuntimeError("drh_server_create failed")
        self._n_out = {}

    def submit(self, r):
        i = C.c_int64()
        if self._lib.drh_server_submit(self._h, C.byref(_to_struct(r)), C.byref(i)) != 0:
            raise RuntimeError("drh_server_submit failed (invalid request)")
        self._n_out[i.value] = n_outputs(r)
        return i.value

    def wait(self, rid, timeout=600.0):
        n = self._n_out[rid]
        mean, var = np.empty(n), np.empty(n)
        res = _Result()
        err = C.create_string_buffer(512)
        st = self._lib.drh_server_wait(self._h, rid, float(timeout), C.byref(res), mean.ctypes.data_as(
            C.POINTER(C.c_double)), var.ctypes.data_as(C.POINTER(C.c_double)), n, err, 512)
        if st == 1:
            raise TimeoutError(f"request {rid} not done after {timeout}s")
        del self._n_out[rid]
        out = {k: getattr(res, k) for k, _ in _Result._fields_}
        out["error"] = err.value.decode()
        out["mean"], out["var"] = (mean, var) if res.status == 0 else (None, None)
        return out

    def cancel(self, rid):
        """0: dropped while queued; 1: running fit told to stop at its next
        HMC iteration; 2: unknown id or already finished.  A cancelled
        request's ``wait`` returns status DR_ECANCELED (7)."""
        return int(self._lib.drh_server_cancel(self._h, int(rid)))

    def close(self):
        if self._h:
            self._lib.drh_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
