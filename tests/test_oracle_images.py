"""Pins oracle/images.py against literal voxel loops that follow the text of
include/dualres/impl/nifti_manipulation.inl one statement at a time:
get_nonzero_indices_impl (:334-360), get_nonzero_data_impl (:297-314),
get_bounding_box (:210-219), get_nonzero_indices_bounded_by_box (:377-398),
get_bounding_box_nonzero_flat_index_impl (:243-275) and apply_mask_impl
(:113-125).  The loops walk the volume in NIfTI voxel order (x fastest) with
scalar arithmetic only, so a transposed axis, a dropped offset or a wrong
missing-voxel rule in the vectorised oracle fails here."""
import math

import numpy as np
import pytest

from oracle import images


def _vol(shape=(5, 6, 7), seed=0, frac_missing=0.35):
    """(nz, ny, nx) volume with zeros and NaNs scattered in, and an empty rim
    on some faces so the bounding box is a strict sub-box."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(shape)
    v[rng.random(shape) < frac_missing] = 0.0
    v[rng.random(shape) < 0.1] = np.nan
    v[0, :, :] = 0.0
    v[:, :, -1] = np.nan
    v[:, 0, :] = 0.0
    return v


def _missing(x):
    return math.isnan(x) or x == 0


def _loop_nonzero_indices(vol):
    nz, ny, nx = vol.shape
    flat = vol.reshape(-1)
    idx = []
    for i in range(nx * ny * nz):
        if not _missing(flat[i]):
            idx.append(i)
    out = []
    for i in idx:
        out.append((i % nx, (i // nx) % ny, (i // (nx * ny)) % nz))
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def _loop_nonzero_data(vol, cast):
    out = []
    for x in vol.reshape(-1):
        y = cast(x)
        if not _missing(y):
            out.append(y)
    return out


def _loop_bounding_box(vol):
    ijk = _loop_nonzero_indices(vol)
    lo = [min(r[a] for r in ijk) for a in range(3)]
    hi = [max(r[a] for r in ijk) for a in range(3)]
    return lo, hi, len(ijk)


def _loop_bounded_by_box(vol, lo, hi):
    out = []
    for r in _loop_nonzero_indices(vol):
        if all(lo[a] <= r[a] <= hi[a] for a in range(3)):
            out.append([r[a] - lo[a] for a in range(3)])
    return np.array(out, dtype=np.int64).reshape(-1, 3)


def _loop_bbox_flat_index(vol):
    lo, hi, _ = _loop_bounding_box(vol)
    dims = [hi[a] - lo[a] + 1 for a in range(3)]
    nz, ny, nx = vol.shape
    flat = vol.reshape(-1)
    index = [None] * (dims[0] * dims[1] * dims[2])
    count = 0
    for k in range(dims[2]):
        for j in range(dims[1]):
            for i in range(dims[0]):
                nii_index = (k + lo[2]) * nx * ny + (j + lo[1]) * nx + (i + lo[0])
                bounded = k * dims[0] * dims[1] + j * dims[0] + i
                if not _missing(flat[nii_index]):
                    index[bounded] = count
                    count += 1
                else:
                    index[bounded] = -1
    return index


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nonzero_indices_and_bounding_box(seed):
    v = _vol(seed=seed)
    assert np.array_equal(images.get_nonzero_indices(v), _loop_nonzero_indices(v))
    lo, hi, n = images.get_bounding_box(v)
    lo2, hi2, n2 = _loop_bounding_box(v)
    assert list(lo) == lo2 and list(hi) == hi2 and n == n2
    # the empty faces make the box strict on x (high), y (low) and z (low)
    assert lo2[1] >= 1 and lo2[2] >= 1 and hi2[0] <= v.shape[2] - 2


@pytest.mark.parametrize("seed", [0, 3])
def test_bounded_by_box_with_a_smaller_box(seed):
    v = _vol(seed=seed)
    lo, hi, _ = _loop_bounding_box(v)
    # the box of the own mask, and a box cut inside it (rows outside dropped)
    for blo, bhi in ((lo, hi), ([lo[0] + 1, lo[1], lo[2] + 1], [hi[0] - 1, hi[1] - 1, hi[2]])):
        got = images.get_nonzero_indices_bounded_by_box(v, np.array(blo), np.array(bhi))
        assert np.array_equal(got, _loop_bounded_by_box(v, blo, bhi))


@pytest.mark.parametrize("seed", [0, 4])
def test_bounding_box_flat_index(seed):
    v = _vol(seed=seed)
    got = images.get_bounding_box_nonzero_flat_index(v)
    want = _loop_bbox_flat_index(v)
    assert got.tolist() == want
    # nonzero order of the box equals the order of get_nonzero_indices
    assert max(want) + 1 == _loop_nonzero_indices(v).shape[0]


def test_nonzero_data_casts_before_the_missing_test():
    """get_nonzero_data_impl casts to ResultType first: a double voxel that
    underflows to 0 in float is dropped from float data but not from double."""
    v = _vol(seed=5)
    v[2, 3, 1] = 1e-50
    v[3, 2, 2] = -1e-60
    d64 = images.get_nonzero_data(v, np.float64)
    d32 = images.get_nonzero_data(v, np.float32)
    w64 = _loop_nonzero_data(v, float)
    w32 = _loop_nonzero_data(v, np.float32)
    assert d64.tolist() == w64
    assert d32.tolist() == [float(x) for x in w32]
    assert len(w64) == len(w32) + 2


def test_apply_mask_loop():
    v = _vol(seed=6)
    rng = np.random.default_rng(7)
    mask = (rng.random(v.shape) < 0.7).astype(np.int32)
    got = images.apply_mask(v, mask)
    want = v.reshape(-1).copy()
    m = mask.reshape(-1)
    for i in range(want.size):
        if m[i] == 0 or math.isnan(want[i]):
            want[i] = 0.0
    assert np.array_equal(got.reshape(-1), want)


def test_voxel_dimensions_column_norms():
    q = np.array([[0.0, -1.8, 0.0, 10.0], [1.8, 0.0, 0.3, -4.0], [0.0, 0.0, 2.3, 7.0], [0, 0, 0, 1]], np.float32)
    got = images.voxel_dimensions(q)
    want = [math.sqrt(sum(float(q[r, c]) ** 2 for r in range(3))) for c in range(3)]
    np.testing.assert_allclose(got, want, rtol=1e-7)
