"""Pins oracle/kriging.py: neighbourhood counts, brute-force pattern, FP64
solve residuals, row normalisation."""
import numpy as np
import pytest

from dualres_amd import problem
from oracle import kriging, images


@pytest.mark.parametrize("vox,count", [(1.0, 1357), (1.5, 437), (2.0, 171)])
def test_neighbourhood_counts(vox, count):
    """Ball of radius 6.9 mm on the grid (integer lattice points inside)."""
    q = problem.qform_axis(vox, (8, 8, 8))
    P = kriging.neighborhood_perturbation(q, 6.9)
    assert P.shape == (count, 3)
    # brute force in float64
    r = int(6.9 / vox)
    g = np.array([[x, y, z] for z in range(-r, r + 1) for y in range(-r, r + 1) for x in range(-r, r + 1)])
    inside = g[np.sqrt(((g * vox) ** 2).sum(1)) <= 6.9]
    assert np.array_equal(P, inside)  # same points, x-fastest order


def _brute_pattern(hr, std, q_hr, q_std, radius):
    """Literal loops of kriging_matrix.h:242-299 (pattern only)."""
    nz, ny, nx = hr.shape
    P = kriging.neighborhood_perturbation(q_hr, radius)
    lo, hi, _ = images.get_bounding_box(hr)
    bd = hi - lo + 1
    bidx = images.get_bounding_box_nonzero_flat_index(hr)
    Mmap = kriging.qform_inverse_map(q_hr, q_std)
    out = []
    for s in images.get_nonzero_indices(std):
        v = np.array([s[0], s[1], s[2], 1.0], np.float32)
        base = [int(np.float32(sum(np.float32(Mmap[r, c] * v[c]) for c in range(4)))) for r in range(3)]
        row = []
        for j, p in enumerate(P):
            h = [base[a] + int(p[a]) for a in range(3)]
            off = h[2] * ny * nx + h[1] * nx + h[0]
            if 0 <= off < hr.size:
                val = hr.ravel()[off]
                if not (np.isnan(val) or val == 0):
                    b = (h[2] - lo[2]) * bd[0] * bd[1] + (h[1] - lo[1]) * bd[0] + (h[0] - lo[0])
                    if 0 <= b < bidx.size and bidx[b] >= 0:
                        row.append((bidx[b], j))
        out.append(row)
    return out


def _problem(seed=1):
    theta = (0.806, 0.132, 1.5)
    p = problem.dual_problem((10, 8, 8), 2.0, theta, seed_gp=seed)
    return p, theta


def test_pattern_brute_force():
    p, theta = _problem()
    radius = 4.6
    rows, P, st = kriging.kriging_pattern(p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    ref = _brute_pattern(p["hr"], p["std"], p["q_hr"], p["q_std"], radius)
    assert len(rows) == len(ref)
    for (col, pidx, kp), r in zip(rows, ref):
        assert [(int(a), int(b)) for a, b in zip(col, pidx)] == r


def test_aligned_base_is_twice_std_index():
    """stdres origin = highres origin + half a highres voxel: (int)(Qh^-1 Qs ijk) = 2 ijk."""
    p, _ = _problem()
    M = kriging.qform_inverse_map(p["q_hr"], p["q_std"])
    for s in images.get_nonzero_indices(p["std"])[:20]:
        v = np.array([s[0], s[1], s[2], 1.0], np.float32)
        mapped = M[:3] @ v
        assert np.array_equal(np.trunc(mapped).astype(int), 2 * s)
        assert np.all(np.abs(mapped - 2 * s - 0.5) < 1e-5)


def test_weights_solve_and_normalise():
    p, theta = _problem(2)
    radius = 4.6
    rp, cols, vals, nrow, ncol, st = kriging.get_sparse_kriging_matrix_data(
        p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    rows, P, _ = kriging.kriging_pattern(p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    K = kriging.perturbation_kernel_distances(P, p["q_hr"], theta).astype(np.float64)
    assert np.all(np.diag(K) == 1.0)
    assert np.allclose(K, K.T)
    for i, (col, pidx, kp) in enumerate(rows):
        w = vals[rp[i]:rp[i + 1]]
        if w.size == 0:
            continue
        assert np.isclose(w.sum(), 1.0, atol=1e-12)
        # K w is proportional to k' (normalisation by the sum)
        Kw = K[np.ix_(pidx, pidx)] @ w
        s = kp.astype(np.float64) @ np.linalg.solve(K[np.ix_(pidx, pidx)], kp.astype(np.float64))
        assert np.linalg.norm(Kw * w.size - kp * (Kw * w.size / kp).mean()) <= 1e-9 * np.linalg.norm(kp) * w.size
        resid = np.linalg.norm(K[np.ix_(pidx, pidx)] @ (w * np.sum(np.linalg.solve(K[np.ix_(pidx, pidx)], kp))) - kp)
        assert resid <= 1e-10 * np.linalg.norm(kp)
        assert s > 0
    assert ncol == int((p["hr"] != 0).sum())
    assert nrow == int((p["std"] != 0).sum())


def test_kernel_values_float_rounding():
    """K and k' entries are float32 roundings of the float64 kernel."""
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 1]])
    q = problem.qform_axis(1.0, (4, 4, 4))
    K = kriging.perturbation_kernel_distances(P, q, (0.9, 0.3, 1.2))
    d = np.sqrt(5.0)
    assert K[0, 2] == np.float32(0.9 * np.exp(-0.3 * d ** 1.2)) or abs(K[0, 2] - 0.9 * np.exp(-np.float32(0.3) * np.float32(d) ** np.float32(1.2))) < 1e-6
