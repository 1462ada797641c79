"""dr_mce_summary_noalloc (device buffers in, device buffers out, every
intermediate in the library's fixed per-device scratch array) against
dr_mce_summary on the same images: pair counts and distances bit-exact,
covariances within 1e-9 of the image variance (the two paths differ only in
how the FFT twiddle tables are rounded).  Two calls in a row, the second on a
smaller image, so the second reuses scratch that still holds the first
call's contents; then a size beyond the scratch limit is refused."""
import ctypes as C

import numpy as np
import pytest

from dualres_amd import _lib as L
from dualres_amd import problem
from dualres_amd.mce import mce_offsets, mce_summary

pytestmark = pytest.mark.gpu


def _image(dims, seed):
    p = problem.single_problem(dims, 1.5, (1.0, 0.3, 1.2), seed_gp=seed, seed_noise=seed + 1, frac=0.48)
    img = p["hr"].copy()
    img[2, 3, 4] = np.nan
    return img, p["q_hr"]


class _Hip:
    """Device buffers for the test through the HIP runtime the engine library
    already loaded (the test allocates; the entry point under test does not)."""

    def __init__(self):
        L.load()
        self.h = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def close(self):
        self.sync()
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(hip, img, q):
    lib = L.load()
    n = mce_offsets().shape[0]
    d_img = hip.put(np.ascontiguousarray(img, dtype=np.float64))
    d_cov = hip.put(np.full(n, np.nan, np.float32))
    d_dist = hip.put(np.full(n, np.nan, np.float32))
    d_np = hip.put(np.full(n, -1, np.int32))
    dims = (C.c_int * 3)(img.shape[2], img.shape[1], img.shape[0])
    qf = np.ascontiguousarray(q, dtype=np.float32).reshape(16)
    L.check(lib.dr_mce_summary_noalloc(d_img, dims, L.fptr(qf), 0, None, d_cov, d_dist, d_np))
    hip.sync()
    return hip.get(d_cov, n, np.float32), hip.get(d_dist, n, np.float32), hip.get(d_np, n, np.int32)


def test_noalloc_mce_summary_matches_allocating_entry_point(gpu_available):
    hip = _Hip()
    for dims, seed in (((40, 36, 30), 1), ((22, 18, 16), 2)):
        img, q = _image(dims, seed)
        print(f"noalloc mce: image {dims}", flush=True)
        cov, dist, npairs = _noalloc(hip, img, q)
        c_ref, d_ref, n_ref = mce_summary(img, q)
        assert np.array_equal(npairs, n_ref)
        assert np.array_equal(dist, d_ref)
        scale = abs(float(c_ref[0]))
        assert np.max(np.abs(cov.astype(np.float64) - c_ref.astype(np.float64))) <= 1e-9 * scale + 1e-7 * scale
    hip.close()


def test_noalloc_mce_summary_refuses_images_beyond_scratch(gpu_available):
    lib = L.load()
    hip = _Hip()
    n = mce_offsets().shape[0]
    d_img = hip.alloc(64)   # never read: the size check comes first
    out = hip.alloc(4 * n)
    dims = (C.c_int * 3)(128, 160, 128)
    qf = np.eye(4, dtype=np.float32).reshape(16)
    rc = lib.dr_mce_summary_noalloc(d_img, dims, L.fptr(qf), 0, None, out, out, out)
    assert rc == L.DR_EDOMAIN
    assert b"limit" in lib.dr_last_error()
    hip.close()
