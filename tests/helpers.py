"""Shared builders for parity tests: matched oracle / engine chains."""
import numpy as np

from dualres_amd import problem
from oracle import kriging, mrp


def small_single(dims=(12, 10, 8), vox=2.0, theta=(0.806, 0.132, 1.0), seed=3):
    return problem.single_problem(dims, vox, theta, seed_gp=seed, seed_noise=seed + 1)


def small_dual(dims=(12, 10, 8), vox=2.0, theta=(0.806, 0.132, 1.5), radius=None, seed=5):
    p = problem.dual_problem(dims, vox, theta, seed_gp=seed, seed_noise=(seed + 1, seed + 2))
    if radius is None:
        radius = 2.3 * vox
    rp, cols, vals, nrow, ncol, st = kriging.get_sparse_kriging_matrix_data(
        p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    import scipy.sparse as sp
    p["W"] = sp.csr_matrix((vals, cols, rp), shape=(nrow, ncol))
    p["radius"] = radius
    return p


def oracle_chain(p, theta, dtype=np.float64, embed=None, seed=11):
    n = 2 if "W" in p else 1
    o = mrp.MultiResParameters(n, theta, p["ijk"], p["q_hr"], rng=mrp.NumpyRng(seed), dtype=dtype, M=embed)
    data = mrp.MultiResData(p["Yh"], p.get("Ys"), p.get("W"), dtype=dtype)
    return o, data


def engine_chain(p, theta, precision="fp64", state_mode="compat", embed=None, seed=11):
    from dualres_amd import Chain
    n = 2 if "W" in p else 1
    c = Chain(p["ijk"], theta, p["q_hr"], n_datasets=n, precision=precision, state_mode=state_mode,
              embed=embed, seed=seed)
    if n == 2:
        c.set_data(p["Yh"], p["Ys"], p["W"])
    else:
        c.set_data(p["Yh"])
    return c


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / max(n, 1e-300)


def random_state(o, rng, imag=True):
    """A prior-like complex state (GP draw / 10 + data) and a momentum draw."""
    N = o.N
    zr, zi = rng.standard_normal(N), rng.standard_normal(N)
    z = (zr + (1j * zi if imag else 0)).reshape(o.shape)
    muh = mrp.F(z) * np.sqrt(o.lam.astype(np.complex128)) / (10 * N)
    muh.reshape(-1)[o.Z] = 0
    mu = mrp.B(muh)
    pr, pi = rng.standard_normal(N), rng.standard_normal(N)
    p = (pr + (1j * pi if imag else 0)).reshape(o.shape)
    pt = mrp.B(p) * np.sqrt(np.abs(1.0 / o.lam.astype(np.complex128) + 1.0)) / N
    pt.reshape(-1)[o.Z] = 0
    p = mrp.F(pt)
    if not imag:
        mu = mu.real + 0j
        p = p.real + 0j
    return mu, p


def dense_posterior(o, Yh, s):
    """Exact posterior of Re(mu) for a single-resolution chain with the noise
    precision s held fixed: precision A = C^-1 + s P^T P, mean A^-1 s P^T Yh,
    with C^-1 applied column by column as mu -> F(B(mu) / (N lambda))
    (the prior term of MultiResParameters.h:337-376).  Needs an empty Z."""
    assert o.Z.size == 0
    N = o.N
    Cinv = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        Cinv[:, j] = mrp.F(mrp.B(e.reshape(o.shape)) / (N * o.lam)).real.reshape(-1)
    A = Cinv.copy()
    A[o.I, o.I] += s
    Sig = np.linalg.inv(A)
    m = Sig[:, o.I] @ (s * np.asarray(Yh))
    C = np.linalg.inv(Cinv)
    return m, Sig, C


def mc_compare(X, mean, var, nb=30):
    """Monte-Carlo check of draws X (iterations x voxels) against exact
    marginal means and variances.  Batch-means standard errors over nb
    batches give per-voxel z-scores of the mean and effective sample sizes;
    the variance ratio's standard error is taken as sqrt(2 / ESS).
    Returns (z, ess, var_ratio, var_z)."""
    X = np.asarray(X)
    K = X.shape[0] - X.shape[0] % nb
    X = X[:K]
    Bm = X.reshape(nb, -1, X.shape[1]).mean(axis=1)
    mcse = Bm.std(axis=0, ddof=1) / np.sqrt(nb)
    z = (X.mean(axis=0) - mean) / mcse
    v = X.var(axis=0)
    ess = v / mcse ** 2
    ratio = v / var
    var_z = (ratio - 1.0) / np.sqrt(2.0 / ess)
    return z, ess, ratio, var_z
