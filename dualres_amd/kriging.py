"""Device kriging-matrix build (kriging_matrix.h:147-343) through the C ABI."""
import ctypes as C

import numpy as np

from . import _lib as L


def kriging_build(hr, q_hr, std, q_std, theta, radius, device=0, return_info=False):
    """Sparse W (n_s x n_h, scipy CSR, float64 values) from masked volumes.

    hr, std: arrays of shape (nz, ny, nx); 0 or NaN marks missing voxels.
    Columns index the highres nonzero voxels in voxel order (Yh order).
    """
    import scipy.sparse as sp

    lib = L.load()
    hr = np.ascontiguousarray(hr, dtype=np.float64)
    std = np.ascontiguousarray(std, dtype=np.float64)
    hd = (C.c_int * 3)(hr.shape[2], hr.shape[1], hr.shape[0])
    sd = (C.c_int * 3)(std.shape[2], std.shape[1], std.shape[0])
    qh = np.ascontiguousarray(q_hr, dtype=np.float32).reshape(16)
    qs = np.ascontiguousarray(q_std, dtype=np.float32).reshape(16)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    h = C.c_void_p()
    L.check(lib.dr_kriging_build(L.dptr(hr), hd, L.fptr(qh), L.dptr(std), sd, L.fptr(qs), L.dptr(th),
                                 float(radius), int(device), C.byref(h)))
    try:
        nrow, ncol, nnz, nu = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        L.check(lib.dr_csr_info(h, C.byref(nrow), C.byref(ncol), C.byref(nnz), C.byref(nu)))
        rp = np.empty(nrow.value + 1, dtype=np.int64)
        cols = np.empty(nnz.value, dtype=np.int32)
        vals = np.empty(nnz.value, dtype=np.float64)
        L.check(lib.dr_csr_copy(h, L.iptr64(rp), L.iptr32(cols), L.dptr(vals)))
        nch, nlu = C.c_int64(), C.c_int64()
        L.check(lib.dr_csr_solver_counts(h, C.byref(nch), C.byref(nlu)))
    finally:
        lib.dr_csr_free(h)
    W = sp.csr_matrix((vals, cols, rp), shape=(nrow.value, ncol.value))
    if return_info:
        return W, {"n_unique_solves": nu.value, "nnz": nnz.value, "n_cholesky": nch.value, "n_lu": nlu.value}
    return W
