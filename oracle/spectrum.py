"""Embedding rule, torus base, kernels and eigenvalues (oracle; test only).

Pinned by: the DFT definition (numpy/scipy pocketfft cross-check), the
invariants sum(lambda)/N = c_0, lambda_0 = sum(c), lambda even and real,
a brute-force O(N^2) DFT on tiny tori, and the separable closed form for
nu = 2 with an axis-aligned qform (tests/test_oracle_spectrum.py).
The float32 distance is computed in a fixed order (row dot products left to
right, then (x^2 + y^2) + z^2); Eigen's own summation order for the 4x4
float product is parity unpinned at the 1-ulp level.
"""
import math

import numpy as np
import scipy.fft as sfft


def rbf(distance, bandwidth, exponent, variance):
    """kernels::rbf  (include/dualres/impl/kernels.inl:7-15)."""
    return variance * np.exp(-bandwidth * np.power(np.abs(distance), exponent))


def rbf_inverse(rho, bandwidth, exponent, variance=1.0):
    """kernels::rbf_inverse  (include/dualres/impl/kernels.inl:19-29)."""
    if rho <= 0 or rho >= 1:
        raise ValueError("rbf_inverse: inverse kernel only for rho between (0, 1)")
    return math.pow(-math.log(rho / variance) / bandwidth, 1.0 / exponent)


def rbf_bandwidth_to_fwhm(bandwidth, exponent):
    """kernels::rbf_bandwidth_to_fwhm  (include/dualres/impl/kernels.inl:34-40)."""
    return 2.0 * math.pow(math.log(2.0) / bandwidth, 1.0 / exponent)


def rational_quadratic(val, psi, nu, variance):
    """kernels::rational_quadratic  (include/dualres/impl/kernels.inl:58-74)."""
    if psi <= 0:
        raise ValueError("rational_quadratic: psi <= 0")
    if nu <= 0:
        raise ValueError("rational_quadratic: nu <= 0")
    if variance <= 0:
        raise ValueError("rational_quadratic: variance <= 0")
    return variance * np.power(1 + psi / nu * val * val, -nu)


def embedding_dims_pow2(max_ijk):
    """lambda grid rule M = 2^ceil(log2(2*max_ijk))  (MultiResParameters.h:191-207)."""
    out = []
    for m in max_ijk:
        d = 2 * int(m)
        if d < 1:
            raise ValueError("circulant_base_3d: Cannot have dimensions <= 1")
        out.append(int(2.0 ** math.ceil(math.log(d) / math.log(2.0))))
    return tuple(out)


def torus_seq(M):
    """Per-axis sequence 0,1,..,M/2,M/2-1,..,1 (impl/circulant_base.inl:60-65)."""
    grid = M // 2 + 1
    seq = np.empty(M, dtype=np.float32)
    i = 0.0
    for ll in range(M):
        if ll < grid:
            i += 1
        else:
            i -= 1
        seq[ll] = i - 1
    return seq


def torus_distance(M, qform):
    """float32 torus distance ||Q*(d(i),d(j),d(k),0)|| on the (M2,M1,M0) grid.

    Restates circulant_base_3d (impl/circulant_base.inl:29-89): the distance
    is evaluated in float and returned as float32; layout is x fastest,
    array shape (M2, M1, M0).
    """
    M0, M1, M2 = M
    Q = np.asarray(qform, dtype=np.float32).reshape(4, 4)
    s0 = torus_seq(M0)[None, None, :]
    s1 = torus_seq(M1)[None, :, None]
    s2 = torus_seq(M2)[:, None, None]
    comps = []
    for r in range(3):
        v = (Q[r, 0] * s0).astype(np.float32)
        v = (v + (Q[r, 1] * s1).astype(np.float32)).astype(np.float32)
        v = (v + (Q[r, 2] * s2).astype(np.float32)).astype(np.float32)
        comps.append(v)
    x2 = (comps[0] * comps[0]).astype(np.float32)
    y2 = (comps[1] * comps[1]).astype(np.float32)
    z2 = (comps[2] * comps[2]).astype(np.float32)
    ss = ((x2 + y2).astype(np.float32) + z2).astype(np.float32)
    return np.sqrt(ss).astype(np.float32)


def eigenvalues(M, qform, theta, covfun="rbf", dtype=np.float64):
    """lambda = F(c) with c = cov(base)  (MultiResParameters.h:232-253).

    theta = (variance tau^2, bandwidth psi, exponent nu) in the reference's
    covariance-parameter order.  Returns the complex eigenvalue array of shape
    (M2, M1, M0) in ``dtype``'s complex type (FFTW forward, unnormalised).
    """
    base = torus_distance(M, qform).astype(dtype)
    tau2, psi, nu = (dtype(t) for t in theta)
    if covfun == "rbf":
        c = rbf(base, psi, nu, tau2).astype(dtype)
    elif covfun == "rq":
        c = rational_quadratic(base, psi, nu, tau2).astype(dtype)
    else:
        raise ValueError("Unknown covariance function")
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    return sfft.fftn(c.astype(cdt), axes=(0, 1, 2))


def nonpositive(lam):
    """Indices (flat, x fastest) with Re(lambda) <= 0  (MultiResParameters.h:258-262)."""
    return np.flatnonzero(lam.real.ravel() <= 0)


def check_nonpositive_count(lam):
    """Throw when too many eigenvalues are nonpositive (MultiResParameters.h:283-284)."""
    nz = nonpositive(lam)
    if nz.size >= lam.size // 2:
        raise ValueError("Too many negative eigen values")
    return nz
