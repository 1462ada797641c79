"""Index and mask helpers on dense volumes (oracle; test only).

Restates include/dualres/impl/nifti_manipulation.inl on plain numpy
volumes of shape (nz, ny, nx) (x fastest, the NIfTI voxel order).  A voxel
is "missing" when it is 0 or NaN.  Pinned by brute-force loops in
tests/test_oracle_images.py.
"""
import numpy as np


def _present(vol):
    v = np.asarray(vol)
    return ~(np.isnan(v) | (v == 0))


def get_nonzero_indices(vol):
    """ijk (n x 3, column-major voxel order)  (nifti_manipulation.inl:334-360)."""
    nz, ny, nx = vol.shape
    flat = np.flatnonzero(_present(vol).ravel())
    return np.stack([flat % nx, (flat // nx) % ny, (flat // (nx * ny)) % nz], axis=1).astype(np.int64)


def get_nonzero_data(vol, dtype=np.float64):
    """Nonzero values in voxel order  (nifti_manipulation.inl:297-314)."""
    v = np.asarray(vol).ravel().astype(dtype)
    return v[_present(v)]


def get_bounding_box(vol):
    """(ijk_min, ijk_max, nnz)  (nifti_manipulation.inl:210-219)."""
    ijk = get_nonzero_indices(vol)
    return ijk.min(axis=0), ijk.max(axis=0), ijk.shape[0]


def get_nonzero_indices_bounded_by_box(vol, bbox_min, bbox_max):
    """ijk relative to a bounding box, rows inside the box  (nifti_manipulation.inl:377-398)."""
    ijk = get_nonzero_indices(vol)
    ok = np.all((ijk >= bbox_min) & (ijk <= bbox_max), axis=1)
    return ijk[ok] - np.asarray(bbox_min)


def get_bounding_box_nonzero_flat_index(vol):
    """Index of each bbox voxel in nonzero order, -1 if missing  (nifti_manipulation.inl:243-275)."""
    lo, hi, _ = get_bounding_box(vol)
    sub = np.asarray(vol)[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
    pres = _present(sub).ravel()
    index = np.full(pres.size, -1, dtype=np.int64)
    index[pres] = np.arange(int(pres.sum()))
    return index


def apply_mask(vol, mask):
    """Zero voxels where mask == 0 or the image is NaN  (nifti_manipulation.inl:113-125)."""
    out = np.array(vol, copy=True)
    out[(np.asarray(mask) == 0) | np.isnan(out)] = 0
    return out


def voxel_dimensions(qform):
    """Column norms of the 3x3 block  (nifti_manipulation.inl:600-605)."""
    Q = np.asarray(qform, dtype=np.float32).reshape(4, 4)
    return np.sqrt((Q[:, :3].astype(np.float32) ** 2).sum(axis=0, dtype=np.float32)).astype(np.float32)
