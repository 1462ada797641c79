"""Pins oracle/spectrum.py: embedding rule, torus base, eigenvalues.

Checks are chosen so that a plausible mistake fails one of them: a wrong
torus sequence (odd/even, off by one), a transposed axis order, a missing
kernel parameter, a normalised FFT, or a wrong sign convention.
"""
import numpy as np
import pytest

from dualres_amd import problem
from oracle import spectrum


def test_embedding_rule_known_answers():
    # MultiResParameters.h:191-207 on the configuration masks
    assert spectrum.embedding_dims_pow2((35, 43, 35)) == (128, 128, 128)
    assert spectrum.embedding_dims_pow2((115, 143, 115)) == (256, 512, 256)
    assert spectrum.embedding_dims_pow2((1, 2, 4)) == (2, 4, 8)
    with pytest.raises(ValueError):
        spectrum.embedding_dims_pow2((0, 3, 3))


def test_torus_sequence():
    assert spectrum.torus_seq(8).tolist() == [0, 1, 2, 3, 4, 3, 2, 1]
    assert spectrum.torus_seq(2).tolist() == [0, 1]
    s = spectrum.torus_seq(12)
    t = np.arange(12)
    assert np.array_equal(s, np.minimum(t, 12 - t))


def test_distance_axis_order_and_values():
    q = problem.qform_axis((1.0, 2.0, 3.0), (4, 4, 4))
    d = spectrum.torus_distance((4, 6, 8), q)
    assert d.shape == (8, 6, 4)  # (M2, M1, M0), x fastest
    assert d[0, 0, 1] == np.float32(1.0)
    assert d[0, 1, 0] == np.float32(2.0)
    assert d[1, 0, 0] == np.float32(3.0)
    assert d[7, 5, 3] == np.float32(np.sqrt(1 + 4 + 9))


def _brute_dft(c):
    M2, M1, M0 = c.shape
    z, y, x = np.meshgrid(np.arange(M2), np.arange(M1), np.arange(M0), indexing="ij")
    out = np.zeros(c.shape, dtype=np.complex128)
    for kz in range(M2):
        for ky in range(M1):
            for kx in range(M0):
                ph = np.exp(-2j * np.pi * (kx * x / M0 + ky * y / M1 + kz * z / M2))
                out[kz, ky, kx] = np.sum(c * ph)
    return out


def test_eigenvalues_brute_force_dft():
    M = (4, 6, 4)
    q = problem.qform_axis(1.5, (4, 4, 4))
    th = (0.8, 0.3, 1.4)
    lam = spectrum.eigenvalues(M, q, th)
    c = spectrum.rbf(spectrum.torus_distance(M, q).astype(np.float64), th[1], th[2], th[0])
    ref = _brute_dft(c)
    assert np.max(np.abs(lam - ref)) < 1e-12 * np.max(np.abs(ref))


def test_eigenvalue_invariants():
    M = (16, 12, 10)
    q = problem.qform_axis(2.0, (8, 6, 5))
    th = (0.806, 0.132, 1.5)
    lam = spectrum.eigenvalues(M, q, th)
    c = spectrum.rbf(spectrum.torus_distance(M, q).astype(np.float64), th[1], th[2], th[0])
    assert np.max(np.abs(lam.imag)) < 1e-12 * np.abs(lam).max()
    # even: lambda(k) == lambda(-k)
    flip = np.roll(lam[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))
    assert np.allclose(lam, flip, rtol=0, atol=1e-12 * np.abs(lam).max())
    assert np.isclose(lam.real.mean(), th[0], rtol=1e-13)  # c_0 = tau^2
    assert np.isclose(lam.real.ravel()[0], c.sum(), rtol=1e-13)


def test_separable_gaussian_closed_form():
    """nu = 2, axis-aligned Q: c is a product of 1D Gaussians, lambda a product of 1D DFTs."""
    M = (12, 10, 8)
    vox = (1.0, 1.5, 2.0)
    q = problem.qform_axis(vox, (6, 5, 4))
    th = (1.3, 0.2, 2.0)
    lam = spectrum.eigenvalues(M, q, th).real
    f = []
    for a in range(3):
        t = np.arange(M[a])
        d = np.minimum(t, M[a] - t) * vox[a]
        f.append(np.fft.fft(np.exp(-th[1] * d * d)).real)
    ref = th[0] * f[2][:, None, None] * f[1][None, :, None] * f[0][None, None, :]
    # float32 distances perturb each value by ~1e-7 relative
    assert np.max(np.abs(lam - ref)) < 1e-6 * np.abs(ref).max()


def test_rational_quadratic_and_inverse():
    assert spectrum.rational_quadratic(0.0, 0.5, 1.0, 2.0) == 2.0
    assert np.isclose(spectrum.rational_quadratic(2.0, 0.5, 1.0, 1.0), 1.0 / 3.0)
    r = spectrum.rbf_inverse(0.1, 0.132, 1.5, 0.806)
    assert np.isclose(spectrum.rbf(r, 0.132, 1.5, 0.806), 0.1, rtol=1e-12)
    with pytest.raises(ValueError):
        spectrum.rbf_inverse(1.0, 0.1, 1.0)
    with pytest.raises(ValueError):
        spectrum.rational_quadratic(1.0, 0.0, 1.0, 1.0)


def test_nonpositive_count_guard():
    M = (8, 4, 4)
    q = problem.qform_axis(1.0, (4, 2, 2))
    lam = spectrum.eigenvalues(M, q, (1.0, 0.0, 1.0))  # constant kernel
    with pytest.raises(ValueError):
        spectrum.check_nonpositive_count(lam)
