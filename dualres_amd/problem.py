"""Synthetic fit problems: ellipsoid masks, seeded GP draws, qforms, indices.

Volumes are numpy arrays of shape (nz, ny, nx) (x fastest, NIfTI voxel
order).  Index conventions follow include/dualres/impl/nifti_manipulation.inl
(nonzero = not 0 and not NaN; ijk in column-major voxel order; data indices
relative to the input-mask bounding box, as fit_model passes them,
gaussian_process_model.h:104-113).
"""
import math

import numpy as np


def qform_axis(vox, dims, origin=None):
    """Axis-aligned qto_xyz with origin -(n-1)/2 * vox (row-major 4x4 float32)."""
    vox = np.broadcast_to(np.asarray(vox, dtype=np.float64), (3,))
    if origin is None:
        origin = [-(dims[a] - 1) / 2.0 * vox[a] for a in range(3)]
    Q = np.eye(4, dtype=np.float32)
    for a in range(3):
        Q[a, a] = vox[a]
        Q[a, 3] = origin[a]
    return Q


def ellipsoid_mask(dims, frac=0.45):
    """Centred ellipsoid, semi-axes frac*dims voxels, centre (n-1)/2; shape (nz, ny, nx)."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = (((x - (nx - 1) / 2) / (frac * nx)) ** 2 + ((y - (ny - 1) / 2) / (frac * ny)) ** 2
         + ((z - (nz - 1) / 2) / (frac * nz)) ** 2)
    return r <= 1.0


def present(vol):
    v = np.asarray(vol)
    return ~(np.isnan(v) | (v == 0))


def nonzero_ijk(vol):
    """(n, 3) int ijk of nonzero voxels in voxel order."""
    nz, ny, nx = vol.shape
    f = np.flatnonzero(present(vol).ravel())
    return np.stack([f % nx, (f // nx) % ny, f // (nx * ny)], axis=1).astype(np.int64)


def bbox(vol):
    ijk = nonzero_ijk(vol)
    return ijk.min(axis=0), ijk.max(axis=0)


def torus_distance(M, vox):
    """Euclidean distance on the torus for an axis-aligned grid (float64)."""
    comps = []
    for a in range(3):
        t = np.arange(M[a])
        comps.append(np.minimum(t, M[a] - t) * float(vox[a]))
    z, y, x = np.meshgrid(comps[2], comps[1], comps[0], indexing="ij")
    return np.sqrt(x * x + y * y + z * z)


def gp_draw(dims, vox, theta, seed, M=None):
    """One zero-mean GP draw on the image grid via circulant embedding (float64).

    Kernel tau2 * exp(-psi * d^nu); torus 2*dims unless given.  Negative
    eigenvalues of the embedding are clipped to 0.
    """
    vox = np.broadcast_to(np.asarray(vox, dtype=np.float64), (3,))
    if M is None:
        M = tuple(2 * int(d) for d in dims)
    tau2, psi, nu = theta
    c = tau2 * np.exp(-psi * np.power(torus_distance(M, vox), nu))
    lam = np.maximum(np.fft.fftn(c).real, 0.0)
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(lam.shape) + 1j * rng.standard_normal(lam.shape)
    f = np.fft.fftn(np.sqrt(lam / lam.size) * e).real
    nx, ny, nz = dims
    return f[:nz, :ny, :nx].copy()


def block_means(vol, k=2):
    nz, ny, nx = vol.shape
    v = vol[: nz - nz % k, : ny - ny % k, : nx - nx % k]
    return v.reshape(nz // k, k, ny // k, k, nx // k, k).mean(axis=(1, 3, 5))


def nudge_zeros(v, mask):
    out = np.array(v, copy=True)
    bad = mask & (out == 0)
    out[bad] = 1e-30
    return out


def single_problem(dims, vox, theta, seed_gp=20261202, seed_noise=1, frac=0.45, noise_sd=1.0):
    """Masked highres image, data indices and values for a single-resolution fit."""
    m = ellipsoid_mask(dims, frac)
    f = gp_draw(dims, vox, theta, seed_gp)
    rng = np.random.default_rng(seed_noise)
    y = f + noise_sd * rng.standard_normal(f.shape)
    y = nudge_zeros(y, m)
    hr = np.where(m, y, 0.0)
    lo, hi = bbox(hr)
    ijk_all = nonzero_ijk(hr)
    flat = np.flatnonzero(present(hr).ravel())
    return {
        "hr": hr, "mask": m, "q_hr": qform_axis(vox, dims), "dims": tuple(dims), "vox": vox,
        "ijk": ijk_all - lo, "Yh": hr.ravel()[flat], "bbox": (lo, hi), "truth": f,
    }


def dual_problem(dims, vox, theta, seed_gp=20261202, seed_noise=(1, 2), frac=0.45, noise_h=1.0, noise_s=0.5):
    """Highres + 2x2x2 block-mean stdres images of the same GP draw."""
    p = single_problem(dims, vox, theta, seed_gp, seed_noise[0], frac, noise_h)
    vox = np.broadcast_to(np.asarray(vox, dtype=np.float64), (3,))
    sdims = tuple(d // 2 for d in dims)
    f = p["truth"]
    s = block_means(f, 2)
    rng = np.random.default_rng(seed_noise[1])
    s = s + noise_s * rng.standard_normal(s.shape)
    sm = ellipsoid_mask(sdims, frac)
    s = nudge_zeros(s, sm)
    std = np.where(sm, s, 0.0)
    q_hr = p["q_hr"]
    q_std = qform_axis(2 * vox, sdims, origin=[q_hr[a, 3] + 0.5 * vox[a] for a in range(3)])
    flat = np.flatnonzero(present(std).ravel())
    p.update({"std": std, "smask": sm, "q_std": q_std, "sdims": sdims, "Ys": std.ravel()[flat]})
    return p


def smooth_even(n):
    """Smallest even 2,3,5,7-smooth integer >= n."""
    m = max(2, int(n))
    while True:
        if m % 2 == 0:
            r = m
            for f in (2, 3, 5, 7):
                while r % f == 0:
                    r //= f
            if r == 1:
                return m
        m += 1


def rbf_inverse(rho, psi, nu, tau2=1.0):
    return math.pow(-math.log(rho / tau2) / psi, 1.0 / nu)
