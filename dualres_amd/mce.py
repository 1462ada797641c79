"""Minimum-contrast summary data on the device (estimate_kernel_parameters.h:53-101)."""
import ctypes as C

import numpy as np

from . import _lib as L


def mce_offsets():
    """The reference's 25,348 voxel offsets (perturbation_matrix.h:62-106), (n, 3) int."""
    lib = L.load()
    n = C.c_int64()
    L.check(lib.dr_mce_offsets(C.byref(n), None))
    out = np.empty((n.value, 3), dtype=np.int32)
    L.check(lib.dr_mce_offsets(C.byref(n), L.iptr32(out)))
    return out


def _dims(img):
    return (C.c_int * 3)(img.shape[2], img.shape[1], img.shape[0])


def mce_summary(img, qform, device=0):
    """(covariance f32, distance f32, npairs i32) per offset; img shape (nz, ny, nx)."""
    lib = L.load()
    img = np.ascontiguousarray(img, dtype=np.float64)
    n = mce_offsets().shape[0]
    cov = np.empty(n, np.float32)
    dist = np.empty(n, np.float32)
    npairs = np.empty(n, np.int32)
    q = np.ascontiguousarray(qform, dtype=np.float32).reshape(16)
    L.check(lib.dr_mce_summary(L.dptr(img), _dims(img), L.fptr(q), int(device), L.fptr(cov), L.fptr(dist),
                               L.iptr32(npairs)))
    return cov, dist, npairs


def mce_binned(img, qform, nbins=64, maxdist=20.0, device=0):
    """Pooled covariance in nbins distance bins on [0, maxdist] mm (float64 bin sums)."""
    lib = L.load()
    img = np.ascontiguousarray(img, dtype=np.float64)
    q = np.ascontiguousarray(qform, dtype=np.float32).reshape(16)
    cov = np.empty(nbins)
    dist = np.empty(nbins)
    n = np.empty(nbins, np.int64)
    L.check(lib.dr_mce_binned(L.dptr(img), _dims(img), L.fptr(q), int(device), int(nbins), float(maxdist),
                              L.dptr(cov), L.dptr(dist), L.iptr64(n)))
    return cov, dist, n


def _host_lib():
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdualres_host.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    f32, i32, f64 = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    lib.drh_rbf_fit.argtypes = [f32, f32, i32, C.c_int64, f64, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double]
    lib.drh_rbf_fit.restype = C.c_int
    return lib


def rbf_fit(cov, dist, npairs, theta0=None, constrained=False, maxdist=1e6, xtol=1e-5):
    """compute_rbf_parameters (estimate_kernel_parameters.h:261-387) through
    the host library: the weighted minimum-contrast fit of (tau^2, psi, nu),
    started where estimate_rbf starts (0.8 cov[0], 0.6, 1.5).  Returns theta
    and the status (0 ok, 1 fit failure, 2 constraint violated)."""
    cov = np.ascontiguousarray(cov, dtype=np.float32)
    dist = np.ascontiguousarray(dist, dtype=np.float32)
    npairs = np.ascontiguousarray(np.minimum(npairs, np.iinfo(np.int32).max), dtype=np.int32)
    th = np.array(theta0 if theta0 is not None else (0.8 * float(cov[0]), 0.6, 1.5), dtype=np.float64)
    rc = _host_lib().drh_rbf_fit(cov.ctypes.data_as(C.POINTER(C.c_float)), dist.ctypes.data_as(C.POINTER(C.c_float)),
                                 npairs.ctypes.data_as(C.POINTER(C.c_int32)), cov.size,
                                 th.ctypes.data_as(C.POINTER(C.c_double)), int(constrained), float(maxdist), -1.0,
                                 -1.0, -1.0, float(xtol))
    return th, rc
