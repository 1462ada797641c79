"""GPU minimum-contrast summary vs the oracle's direct sums.

Pair counts bit-exact; covariances within 1e-9 of the variance (the FFT
correlation and the direct FP64 sums differ only by rounding); distances
bit-exact (same float32 evaluation order)."""
import numpy as np
import pytest

from dualres_amd import problem
from dualres_amd.mce import mce_binned, mce_offsets, mce_summary
from oracle import mce

pytestmark = pytest.mark.gpu


def _image(dims=(22, 18, 16), seed=0):
    theta = (1.0, 0.3, 1.2)
    p = problem.single_problem(dims, 1.5, theta, seed_gp=seed, seed_noise=seed + 1, frac=0.48)
    img = p["hr"].copy()
    img[3, 4, 5] = np.nan
    return img, p["q_hr"]


def test_offsets_match_oracle(gpu_available):
    assert np.array_equal(mce_offsets(), mce.perturbation_matrix_3d())


def test_mce_summary_matches_direct_sums(gpu_available):
    img, q = _image()
    cov, dist, n = mce_summary(img, q)
    c_o, d_o, n_o = mce.compute_mce_summary_data(img, q)
    assert np.array_equal(n, n_o)
    assert np.array_equal(dist, d_o)
    assert np.max(np.abs(cov.astype(np.float64) - c_o)) <= 1e-9 * abs(float(c_o[0])) + 1e-6 * np.abs(c_o).max() * 1e-3


def test_mce_binned_pools_offsets(gpu_available):
    img, q = _image(seed=3)
    nb, md = 8, 6.0
    cov, dist, n = mce_binned(img, q, nbins=nb, maxdist=md)
    # pool the oracle's direct sums over the same half-space offsets
    r = int(np.ceil(md / 1.5))
    P = np.array([[x, y, z] for z in range(0, r + 1) for y in range(-r, r + 1) for x in range(-r, r + 1)
                  if not (z == 0 and (y < 0 or (y == 0 and x < 0)))])
    d = mce.offset_distances(P, q)
    keep = d <= np.float32(md)
    P, d = P[keep], d[keep]
    s = mce.mce_sums(img, P)
    b = np.minimum((d / (md / nb)).astype(int), nb - 1)
    for k in range(nb):
        sel = b == k
        nn = s[sel, 3].sum()
        assert n[k] == nn
        if nn > 1:
            ref = (s[sel, 0].sum() - s[sel, 1].sum() * s[sel, 2].sum() / nn) / (nn - 1)
            assert abs(cov[k] - ref) <= 1e-9 * abs(cov[0]) + 1e-12


def test_mce_full_size_closed_forms(gpu_available):
    """At the estimation batch's full image size, 128x160x128: with every
    voxel present the pair count of offset p is prod_a (n_a - |p_a|) for all
    25,348 offsets (exact), a constant image has covariance 0 at every offset,
    and the 64 bins to 20 mm pool exactly the closed-form counts of their
    half-space offsets."""
    dims = (128, 160, 128)
    q = problem.qform_axis(1.0, dims)
    rng = np.random.default_rng(2)
    img = rng.standard_normal(dims[::-1]) + 5.0
    P = mce_offsets().astype(np.int64)
    want = np.prod(np.array(dims)[None, :] - np.abs(P), axis=1)
    cov, dist, n = mce_summary(img, q)
    assert np.array_equal(n.astype(np.int64), want)
    assert np.array_equal(dist, mce.offset_distances(P, q))
    const = np.full(dims[::-1], 2.5)
    cov_c, _, n_c = mce_summary(const, q)
    assert np.array_equal(n_c.astype(np.int64), want)
    assert np.max(np.abs(cov_c)) < 1e-9 * 2.5 ** 2
    nb, md = 64, 20.0
    bc, bd, bn = mce_binned(const, q, nbins=nb, maxdist=md)
    r = 20
    H = np.array([[x, y, z] for z in range(0, r + 1) for y in range(-r, r + 1) for x in range(-r, r + 1)
                  if not (z == 0 and (y < 0 or (y == 0 and x < 0)))])
    d = mce.offset_distances(H, q)
    keep = d <= np.float32(md)
    H, d = H[keep], d[keep]
    b = np.minimum((d / (md / nb)).astype(int), nb - 1)
    cnt = np.prod(np.array(dims)[None, :] - np.abs(H), axis=1)
    for k in range(nb):
        assert bn[k] == int(cnt[b == k].sum()), k
    assert np.max(np.abs(bc[bn > 1])) < 1e-9 * 2.5 ** 2
