"""Pins oracle/mce.py: offset set, direct sums, closed-form pair counts,
and the LSQ objective/weights."""
import numpy as np

from dualres_amd import problem
from oracle import mce


def test_offset_set():
    P = mce.perturbation_matrix_3d()
    assert P.shape == (25348, 3)
    assert P[0].tolist() == [0, 0, 0]
    assert P[1].tolist() == [1, 0, 0]
    assert len({tuple(r) for r in P.tolist()}) == P.shape[0]
    # raster offsets 19..25 along each axis are present
    s = {tuple(r) for r in P.tolist()}
    for k in range(19, 26):
        assert (k, 0, 0) in s and (0, k, 0) in s and (0, 0, k) in s


def _direct(vol, P):
    nz, ny, nx = vol.shape
    out = np.zeros((P.shape[0], 4))
    for i, (px, py, pz) in enumerate(P):
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    a = vol[z, y, x]
                    if np.isnan(a) or a == 0:
                        continue
                    X, Y, Z = x + px, y + py, z + pz
                    if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz:
                        b = vol[Z, Y, X]
                        if not (np.isnan(b) or b == 0):
                            out[i] += [a * b, a, b, 1]
    return out


def test_direct_sums_small():
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((5, 6, 7))
    vol[rng.random(vol.shape) < 0.2] = 0
    vol[1, 2, 3] = np.nan
    P = mce.perturbation_matrix_3d()[:60]
    assert np.allclose(mce.mce_sums(vol, P), _direct(vol, P), rtol=1e-12, atol=1e-12)


def test_constant_image_and_full_box_counts():
    vol = np.full((6, 7, 8), 2.5)
    q = problem.qform_axis(1.0, (8, 7, 6))
    P = mce.perturbation_matrix_3d()
    cov, dist, n = mce.compute_mce_summary_data(vol, q, P)
    exp_n = np.prod(np.maximum(np.array([8, 7, 6])[None, :] - np.abs(P), 0), axis=1)
    assert np.array_equal(n, exp_n)
    assert np.all(np.abs(cov[n > 1]) < 1e-12)
    assert np.allclose(dist, np.sqrt((P.astype(float) ** 2).sum(1)), rtol=1e-6)


def test_lsq_weights_and_objective():
    cov = np.array([1.0, 0.5, 0.5, 0.25, 0.7], np.float32)
    dist = np.array([0.0, 1.0, 1.0, 2.0, 1.5], np.float32)
    n = np.array([10, 5, 5, 3, 0], np.int32)
    x, y, w = mce.lsq_data(cov, dist, n)
    assert x.tolist() == [1.0, 1.0, 2.0]  # offset 0 and empty offsets dropped
    assert np.allclose(w, [1 / np.sqrt(2), 1 / np.sqrt(2), 1.0])
    th = [1.0, np.log(2.0), 1.0]  # rho(1) = 0.5, rho(2) = 0.25
    assert mce.rbf_least_squares(th, x, y, w) < 1e-30
    assert mce.rbf_least_squares([0.5, np.log(2.0), 1.0], x, y, w) > 0
