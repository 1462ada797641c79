"""Minimum-contrast summary data and the RBF least-squares fit (oracle; test only).

Restates include/dualres/perturbation_matrix.h:20-106 and
include/dualres/estimate_kernel_parameters.h:53-151,241-387.

Pinned by: brute-force direct sums on tiny images, a constant image
(covariance 0), the closed-form pair count of a full box, and the unique
offset count 25,348 (tests/test_oracle_mce.py).  The COBYLA iterate
sequence of NLopt is parity unpinned (NLopt is absent); the fit is pinned by
objective optimality against scipy's COBYLA/Nelder-Mead and by recovery of a
synthetic truth.
"""
import numpy as np

U_DIRECTIONS = np.array([
    [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [-1, 1, 0],
    [1, 0, -1], [0, 1, 1], [0, 1, -1], [1, 1, 1], [1, 1, -1], [-1, 1, 1], [-1, 1, -1]],
    dtype=np.int64)  # perturbation_matrix.h:20-38


def perturbation_permutations_3d(n):
    """{1..n}^3, x fastest  (perturbation_matrix.h:45-55)."""
    z, y, x = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int64)


def perturbation_matrix_3d(nvox_angles=18, nvox_raster=25):
    """Unique MCE offsets in first-occurrence order  (perturbation_matrix.h:62-106)."""
    A = perturbation_permutations_3d(nvox_angles)
    P = (A[:, None, :] * U_DIRECTIONS[None, :, :]).reshape(-1, 3)
    ras = np.zeros((3 * (nvox_raster - nvox_angles), 3), dtype=np.int64)
    for i in range(nvox_raster - nvox_angles):
        for j in range(3):
            ras[i * 3 + j, j] = nvox_angles + i + 1
    P = np.concatenate([P, ras], axis=0)
    _, first = np.unique(P, axis=0, return_index=True)
    return P[np.sort(first)]


def offset_distances(P, qform):
    """||[p, 0] Q^T|| in float32  (estimate_kernel_parameters.h:70)."""
    Q = np.asarray(qform, dtype=np.float32).reshape(4, 4)
    Pf = P.astype(np.float32)
    ss = np.zeros(P.shape[0], dtype=np.float32)
    for r in range(4):
        v = (Pf[:, 0] * Q[r, 0]).astype(np.float32)
        v = (v + (Pf[:, 1] * Q[r, 1]).astype(np.float32)).astype(np.float32)
        v = (v + (Pf[:, 2] * Q[r, 2]).astype(np.float32)).astype(np.float32)
        ss = (ss + (v * v).astype(np.float32)).astype(np.float32)
    return np.sqrt(ss).astype(np.float32)


def mce_sums(vol, P):
    """Per-offset (sum_xy, sum_x, sum_y, npairs) in float64 (estimate_kernel_parameters.h:72-93).

    Vectorised over voxels for each offset; the voxel pair (x, x+p) counts
    when both are in bounds and non-missing.
    """
    v = np.asarray(vol, dtype=np.float64)
    nz, ny, nx = v.shape
    pres = ~(np.isnan(v) | (v == 0))
    vz = np.where(pres, v, 0.0)
    out = np.zeros((P.shape[0], 4))
    for i, (px, py, pz) in enumerate(P):
        xs = slice(max(0, -px), min(nx, nx - px))
        ys = slice(max(0, -py), min(ny, ny - py))
        zs = slice(max(0, -pz), min(nz, nz - pz))
        xd = slice(xs.start + px, xs.stop + px)
        yd = slice(ys.start + py, ys.stop + py)
        zd = slice(zs.start + pz, zs.stop + pz)
        if xs.start >= xs.stop or ys.start >= ys.stop or zs.start >= zs.stop:
            continue
        a = vz[zs, ys, xs]
        b = vz[zd, yd, xd]
        m = pres[zs, ys, xs] & pres[zd, yd, xd]
        out[i, 0] = np.sum(a * b * m)
        out[i, 1] = np.sum(a * m)
        out[i, 2] = np.sum(b * m)
        out[i, 3] = np.sum(m)
    return out


def compute_mce_summary_data(vol, qform, P=None):
    """(covariance f32, distance f32, npairs i32)  (estimate_kernel_parameters.h:53-101)."""
    if P is None:
        P = perturbation_matrix_3d()
    s = mce_sums(vol, P)
    n = s[:, 3].astype(np.int64)
    cov = np.zeros(P.shape[0], dtype=np.float32)
    ok = n > 1
    cov[ok] = ((s[ok, 0] - s[ok, 1] * s[ok, 2] / n[ok]) / (n[ok] - 1)).astype(np.float32)
    return cov, offset_distances(P, qform), n.astype(np.int32)


def lsq_data(cov, dist, npairs, maxdist=1e6):
    """(x, y, w) for the weighted LSQ: w = 1/sqrt(multiplicity of x)  (estimate_kernel_parameters.h:317-357)."""
    keep = (np.arange(len(npairs)) >= 1) & (npairs != 0) & (dist <= maxdist)
    x = np.abs(dist[keep].astype(np.float64))
    y = cov[keep].astype(np.float64)
    _, inv, cnt = np.unique(x, return_inverse=True, return_counts=True)
    w = np.sqrt(1.0 / cnt[inv])
    return x, y, w


def rbf_least_squares(theta, x, y, w):
    """Objective of _rbf_least_squares (estimate_kernel_parameters.h:122-151)."""
    rho = np.exp(-theta[1] * np.power(x, theta[2]))
    r = w * (y - theta[0] * rho)
    return float(np.sum(r * r / x.size))


def bounds(cov0, variance=-1.0, bandwidth=-1.0, exponent=-1.0, eps0=1e-5):
    """Box bounds and fixed parameters (estimate_kernel_parameters.h:272-298)."""
    lb = [eps0, eps0, eps0]
    ub = [float(cov0), np.inf, 2.0 - eps0]
    if variance - eps0 > 0:
        lb[0] = min(variance - eps0, max(cov0 - eps0, eps0))
        ub[0] = min(variance + eps0, cov0)
    if exponent - eps0 > 0 and exponent + eps0 <= 2.0:
        lb[2] = exponent - eps0
        ub[2] = exponent + eps0
    if bandwidth - eps0 > 0:
        lb[1] = bandwidth - eps0
        ub[1] = bandwidth + eps0
    return lb, ub
