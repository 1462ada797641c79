"""Sparse local-kriging matrix W (oracle; test only).

Restates include/dualres/kriging_matrix.h:41-343 on numpy volumes of shape
(nz, ny, nx).  The sparsity pattern (which high-resolution neighbours enter
each row, and in which order) follows the reference exactly, including its
flat-offset range check (kriging_matrix.h:263-266); a bounding-box index that
falls outside the bounded index table (undefined behaviour in the reference,
kriging_matrix.h:275-284) is skipped and counted in ``stats['oob']``.

Kernel values in K and k' are float32 (as in the reference, which evaluates
rbf<float>); they are computed here as the float32 rounding of the float64
kernel, so they do not depend on a platform's float exp/pow ulps.  Weights
are solved in float64 (the reference solves in float32 with Eigen's
colPivHouseholderQr, kriging_matrix.h:309-313; those float32 bits are parity
unpinned).  Pinned by: brute-force integer pattern enumeration and the
FP64 residual ||K w - k'|| / ||k'|| (tests/test_oracle_kriging.py).
"""
import numpy as np

from . import images, spectrum


def expand_grid_3d(r):
    """Integer offsets in [-r, r]^3, x fastest  (kriging_matrix.h:41-56)."""
    r = [int(v) for v in r]
    if min(r) < 0:
        raise ValueError("expand_grid_3d: pos_griddims must all be >= 0")
    z, y, x = np.meshgrid(np.arange(-r[2], r[2] + 1), np.arange(-r[1], r[1] + 1),
                          np.arange(-r[0], r[0] + 1), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int64)


def _mm(P, qform):
    """float32 rows of [P, 0] * Q^T, accumulated left to right."""
    Q = np.asarray(qform, dtype=np.float32).reshape(4, 4)
    Pf = P.astype(np.float32)
    out = np.empty((P.shape[0], 3), dtype=np.float32)
    for r in range(3):
        v = (Pf[:, 0] * Q[r, 0]).astype(np.float32)
        v = (v + (Pf[:, 1] * Q[r, 1]).astype(np.float32)).astype(np.float32)
        v = (v + (Pf[:, 2] * Q[r, 2]).astype(np.float32)).astype(np.float32)
        out[:, r] = v
    return out


def _norm3(v):
    v = v.astype(np.float32)
    s = ((v[..., 0] * v[..., 0]).astype(np.float32) + (v[..., 1] * v[..., 1]).astype(np.float32)).astype(np.float32)
    s = (s + (v[..., 2] * v[..., 2]).astype(np.float32)).astype(np.float32)
    return np.sqrt(s).astype(np.float32)


def neighborhood_perturbation(qform, radius):
    """Offsets within ``radius`` mm  (kriging_matrix.h:63-106)."""
    if radius < 0:
        raise ValueError("neighborhood_perturbation: radius cannot be negative")
    vd = images.voxel_dimensions(qform)
    rng = (np.float32(radius) / vd).astype(np.float32).astype(np.int64)
    P = expand_grid_3d(rng)
    keep = _norm3(_mm(P, qform)) <= np.float32(radius)
    if keep.sum() == 0:
        raise ValueError("neighborhood_perturbation: no perturbations within radius")
    return P[keep]


def perturbation_kernel_distances(P, qform, theta, dtype=np.float32):
    """K(i, j) = rbf(||x_i - x_j||), unit diagonal  (kriging_matrix.h:111-138)."""
    X = _mm(P, qform)
    D = _norm3(X[:, None, :] - X[None, :, :]).astype(np.float64)
    t = [float(np.float32(v)) for v in theta]
    K = spectrum.rbf(D, t[1], t[2], t[0]).astype(np.float32).astype(dtype)
    np.fill_diagonal(K, 1)
    return K


def qform_inverse_map(q_hr, q_std):
    """Qh^{-1} * Qs in float32 (kriging_matrix.h:193-194)."""
    Qh = np.asarray(q_hr, dtype=np.float32).reshape(4, 4)
    Qs = np.asarray(q_std, dtype=np.float32).reshape(4, 4)
    return (np.linalg.inv(Qh.astype(np.float64)) @ Qs.astype(np.float64)).astype(np.float32)


def kriging_pattern(hr, std, q_hr, q_std, theta, radius):
    """Per-row neighbour lists: (cols, pidx, kprime) for every stdres nonzero voxel.

    Restates the loop at kriging_matrix.h:242-299.  Returns lists of arrays
    plus P and a stats dict.
    """
    hr = np.asarray(hr)
    nz, ny, nx = hr.shape
    nvox = hr.size
    flat_hr = hr.ravel()
    P = neighborhood_perturbation(q_hr, np.float32(radius))
    lo, hi, _ = images.get_bounding_box(hr)
    bdims = hi - lo + 1
    bidx = images.get_bounding_box_nonzero_flat_index(hr)
    Mmap = qform_inverse_map(q_hr, q_std)
    Qh = np.asarray(q_hr, dtype=np.float32).reshape(4, 4)
    Qs = np.asarray(q_std, dtype=np.float32).reshape(4, 4)
    ijk_std = images.get_nonzero_indices(std)
    tf = [float(np.float32(v)) for v in theta]
    eps0 = np.float32(1e-5)
    rows = []
    oob = 0
    for i in range(ijk_std.shape[0]):
        s4 = np.array([ijk_std[i, 0], ijk_std[i, 1], ijk_std[i, 2], 1], dtype=np.float32)
        base = np.empty(4, dtype=np.float32)
        for r in range(4):
            acc = np.float32(0)
            for c in range(4):
                acc = np.float32(acc + np.float32(Mmap[r, c] * s4[c]))
            base[r] = acc
        base = np.trunc(base).astype(np.int64)[:3]
        hr_ijk = base[None, :] + P
        off = hr_ijk[:, 2] * ny * nx + hr_ijk[:, 1] * nx + hr_ijk[:, 0]
        ok = (off >= 0) & (off < nvox)
        val = np.zeros(P.shape[0])
        val[ok] = flat_hr[off[ok]]
        ok &= ~(np.isnan(val) | (val == 0))
        b = ((hr_ijk[:, 2] - lo[2]) * bdims[0] * bdims[1] + (hr_ijk[:, 1] - lo[1]) * bdims[0]
             + (hr_ijk[:, 0] - lo[0]))
        inb = (b >= 0) & (b < bidx.size)
        oob += int((ok & ~inb).sum())
        ok &= inb
        col = np.full(P.shape[0], -1, dtype=np.int64)
        col[ok] = bidx[b[ok]]
        ok &= col >= 0
        sel = np.flatnonzero(ok)
        # distance ||Qs*(ijk_s,1) - Qh*(ijk_h,1)|| + 1e-5 in float32
        xs = np.empty(4, dtype=np.float32)
        for r in range(4):
            acc = np.float32(0)
            for c in range(4):
                acc = np.float32(acc + np.float32(Qs[r, c] * s4[c]))
            xs[r] = acc
        h4 = np.concatenate([hr_ijk[sel], np.ones((sel.size, 1), np.int64)], axis=1).astype(np.float32)
        xh = np.zeros((sel.size, 4), dtype=np.float32)
        for r in range(4):
            acc = np.zeros(sel.size, dtype=np.float32)
            for c in range(4):
                acc = (acc + (h4[:, c] * Qh[r, c]).astype(np.float32)).astype(np.float32)
            xh[:, r] = acc
        d = (xs[None, :] - xh).astype(np.float32)
        ss = np.zeros(sel.size, dtype=np.float32)
        for c in range(4):
            ss = (ss + (d[:, c] * d[:, c]).astype(np.float32)).astype(np.float32)
        dist = (np.sqrt(ss).astype(np.float32) + eps0).astype(np.float32)
        kp = spectrum.rbf(dist.astype(np.float64), tf[1], tf[2], tf[0]).astype(np.float32)
        rows.append((col[sel], sel, kp))
    return rows, P, {"oob": oob, "n_rows": ijk_std.shape[0], "ncol": int((bidx >= 0).sum())}


def get_sparse_kriging_matrix_data(hr, std, q_hr, q_std, theta, radius):
    """CSR (row_ptr, cols, vals, nrow, ncol) with float64 weights (kriging_matrix.h:147-343)."""
    rows, P, st = kriging_pattern(hr, std, q_hr, q_std, theta, radius)
    K = perturbation_kernel_distances(P, q_hr, theta, dtype=np.float32).astype(np.float64)
    row_ptr = [0]
    cols, vals = [], []
    empty = 0
    for col, pidx, kp in rows:
        if col.size == 0:
            empty += 1
            row_ptr.append(row_ptr[-1])
            continue
        w = np.linalg.solve(K[np.ix_(pidx, pidx)], kp.astype(np.float64))
        s = w.sum()
        if s == 0:
            s = 1.0
        if np.isnan(s):
            row_ptr.append(row_ptr[-1])
            continue
        w = w / s
        cols.append(col)
        vals.append(w)
        row_ptr.append(row_ptr[-1] + col.size)
    cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    vals = np.concatenate(vals) if vals else np.zeros(0)
    st["empty"] = empty
    return np.asarray(row_ptr, np.int64), cols, vals, st["n_rows"], st["ncol"], st
