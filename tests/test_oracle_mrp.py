"""Pins oracle/mrp.py (MultiResParameters restatement) by dense algebra,
finite differences and leapfrog invariants on tiny tori."""
import numpy as np
import pytest
import scipy.sparse as sp

from dualres_amd import problem
from oracle import mrp, spectrum

THETA = (0.806, 1.0, 1.5)  # all eigenvalues positive on the 8x4x4 torus


def _tiny(dual=False, seed=0):
    rng = np.random.default_rng(seed)
    dims = (4, 3, 3)
    ijk = np.array([[i, j, k] for k in range(3) for j in range(3) for i in range(4) if (i + j + k) % 3 != 1])
    q = problem.qform_axis(1.0, dims)
    o = mrp.MultiResParameters(2 if dual else 1, THETA, ijk, q, rng=mrp.NumpyRng(seed))
    Yh = rng.standard_normal(ijk.shape[0])
    if dual:
        W = sp.random(5, ijk.shape[0], density=0.4, random_state=seed, format="csr")
        data = mrp.MultiResData(Yh, rng.standard_normal(5), W)
    else:
        data = mrp.MultiResData(Yh)
    o.sigma_sq_inv = [1.7, 0.6][: o.n_datasets]
    z = rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape)
    o.mu = z
    return o, data


def _dense_C(o):
    M0, M1, M2 = o.M
    base = spectrum.torus_distance(o.M, np.asarray(problem.qform_axis(1.0, (4, 3, 3)))).astype(np.float64)
    c = spectrum.rbf(base, THETA[1], THETA[2], THETA[0]).ravel()
    idx = np.arange(o.N)
    x, y, z = idx % M0, (idx // M0) % M1, idx // (M0 * M1)
    dx = (x[:, None] - x[None, :]) % M0
    dy = (y[:, None] - y[None, :]) % M1
    dz = (z[:, None] - z[None, :]) % M2
    return c[dx + M0 * (dy + M1 * dz)]


@pytest.mark.parametrize("dual", [False, True])
def test_gradient_dense_circulant(dual):
    o, data = _tiny(dual)
    assert o.Z.size == 0
    C = _dense_C(o)
    mu = o.mu.reshape(-1)
    r = o.sigma_sq_inv[0] * (data.Yh - mu[o.I].real)
    if dual:
        r = r + o.sigma_sq_inv[1] * (data.W.T @ (data.Ys - data.W @ mu[o.I].real))
    g_dense = -np.linalg.solve(C, mu)
    g_dense[o.I] += r
    g = o.compute_gradient(data, o.mu).reshape(-1)
    assert np.max(np.abs(g - g_dense)) < 1e-10 * np.max(np.abs(g_dense))


def test_log_prior_dense():
    o, data = _tiny()
    C = _dense_C(o)
    mu = o.mu.reshape(-1)
    lp = -0.5 * (mu.real @ np.linalg.solve(C, mu.real) + mu.imag @ np.linalg.solve(C, mu.imag))
    assert np.isclose(o._log_prior(o.mu), lp, rtol=1e-11)


@pytest.mark.parametrize("dual", [False, True])
def test_gradient_finite_differences(dual):
    o, data = _tiny(dual, seed=3)
    g = o.compute_gradient(data, o.mu).reshape(-1)
    h = 1e-5
    for j in [0, 5, 17, o.I[3], o.I[-1]]:
        e = np.zeros(o.N, dtype=complex)
        e[j] = h
        up = o._log_posterior(data, o.mu + e.reshape(o.shape))
        dn = o._log_posterior(data, o.mu - e.reshape(o.shape))
        assert np.isclose((up - dn) / (2 * h), g[j].real, rtol=1e-6, atol=1e-8)
        ei = 1j * e
        up = o._log_posterior(data, o.mu + ei.reshape(o.shape))
        dn = o._log_posterior(data, o.mu - ei.reshape(o.shape))
        assert np.isclose((up - dn) / (2 * h), g[j].imag, rtol=1e-6, atol=1e-8)


def test_leapfrog_reversible_and_second_order():
    o, data = _tiny(True, seed=5)
    o._compute_mass_matrix_eigen_values()
    rng = np.random.default_rng(1)
    p0 = mrp.F(rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape))
    mu0 = o.mu.copy()
    mu1, p1 = o.leapfrog(data, mu0.copy(), p0.copy(), 0.05, 7)
    mu2, p2 = o.leapfrog(data, mu1.copy(), -p1, 0.05, 7)
    assert np.max(np.abs(mu2 - mu0)) < 1e-10 * np.max(np.abs(mu0))
    assert np.max(np.abs(-p2 - p0)) < 1e-9 * np.max(np.abs(p0))

    def dH(eps):
        L = int(round(0.4 / eps))
        m, p = o.leapfrog(data, mu0.copy(), p0.copy(), eps, L)
        return (o._log_posterior(data, m) + o._potential_energy(p)) - (o._log_posterior(data, mu0) + o._potential_energy(p0))

    r = dH(0.02) / dH(0.01)
    assert 3.0 < r < 5.0


def test_leapfrog_harmonic_closed_form():
    """Every torus point observed, no W: mass matrix = posterior precision, so
    each mode of mu - mu* follows the unit-frequency leapfrog map."""
    o, _ = _tiny(False, seed=7)
    o.I = np.arange(o.N)
    rng = np.random.default_rng(2)
    Y = rng.standard_normal(o.N)
    data = mrp.MultiResData(Y)
    o._compute_mass_matrix_eigen_values()
    lamm = o.lam_mass.reshape(-1)
    prec = (1.0 / o.lam.reshape(-1) + o.sigma_sq_inv[0])
    assert np.allclose(lamm, prec)
    # posterior mean mu* = M^-1 sigma Y (spectral solve)
    mstar = mrp.B(mrp.F(o.sigma_sq_inv[0] * Y.reshape(o.shape)) / prec.reshape(o.shape)) / o.N
    u0 = (rng.standard_normal(o.N)).reshape(o.shape)
    # a real momentum: the data term only drives the real part of the state
    p0 = rng.standard_normal(o.shape) + 0j
    eps, L = 0.3, 5
    mu1, p1 = o.leapfrog(data, (mstar + u0).astype(complex), p0.copy(), eps, L)
    # w = M^-1 p in real space
    w0 = mrp.F(mrp.B(p0) / (o.lam_mass * o.N))
    w1 = mrp.F(mrp.B(p1) / (o.lam_mass * o.N))
    A = np.array([[1 - eps ** 2 / 2, eps], [-eps * (1 - eps ** 2 / 4), 1 - eps ** 2 / 2]])
    AL = np.linalg.matrix_power(A, L)
    u_pred = AL[0, 0] * u0 + AL[0, 1] * w0
    w_pred = AL[1, 0] * u0 + AL[1, 1] * w0
    assert np.max(np.abs((mu1 - mstar) - u_pred)) < 1e-9
    assert np.max(np.abs(w1 - w_pred)) < 1e-9


def test_update_accept_and_energies_consistent():
    o, data = _tiny(True, seed=9)
    rng = np.random.default_rng(4)
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)], uniforms=[0.0], gammas=[2.0, 3.0])
    mu_start = o.mu.copy()
    lp_start = o._log_posterior(data, mu_start)
    mh = o.update(data, 0.02, 3)
    assert 0.0 <= mh <= 1.0
    assert o.accepted  # u = 0 < R always unless R = 0
    assert o.sigma_sq_inv == [2.0, 3.0]
    # H0 = E0 + log posterior of the start, with E0 = -1/2 sum |z|^2
    e0 = -0.5 * sum(a * a + b * b for a, b in zip(zr.tolist(), zi.tolist()))
    assert o.initial_energy == pytest.approx(e0 + lp_start, rel=1e-12)


def _neg(o, f):
    """Flat index of the mode -k for the flat index f of k."""
    M0, M1, M2 = o.M
    x, y, z = f % M0, (f // M0) % M1, f // (M0 * M1)
    return (-x) % M0 + M0 * ((-y) % M1 + M1 * ((-z) % M2))


def _fresh_momentum_kinetic(o, zr, zi):
    o._compute_mass_matrix_eigen_values()
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)])
    e0 = o._sample_momentum()
    return e0, o._potential_energy(o.momentum)


def test_sample_momentum_kinetic_energy_equals_e0():
    """_sample_momentum (MultiResParameters.h:783-815): p = F(sqrt(lambda_m)
    B(z) / N).  Since B F = N I, B(p) = sqrt(lambda_m) B(z), so the kinetic
    energy of the fresh momentum is -1/2 sum_{k not in Z} |B(z)_k|^2 / N, which
    by Parseval is E0 = -1/2 sum |z|^2 when Z is empty.  A dropped square
    root or a lost 1/N breaks the equality."""
    o, _ = _tiny(True, seed=11)
    assert o.Z.size == 0
    rng = np.random.default_rng(12)
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    e0, k = _fresh_momentum_kinetic(o, zr, zi)
    assert e0 == pytest.approx(-0.5 * float(np.sum(zr ** 2 + zi ** 2)), rel=1e-14)
    assert k == pytest.approx(e0, rel=1e-12)
    # a non-empty low-rank set: the kinetic energy loses exactly the energy of
    # B(z) on Z (Z is closed under k -> -k because lambda is even)
    o.Z = np.array([1, _neg(o, 1), 5, _neg(o, 5)])
    lam = o.lam.reshape(-1).copy()
    assert np.allclose(lam[[1, 5]], lam[[_neg(o, 1), _neg(o, 5)]])
    e0, k = _fresh_momentum_kinetic(o, zr, zi)
    bz = mrp.B((zr + 1j * zi).reshape(o.shape)).reshape(-1)
    keep = np.ones(o.N, bool)
    keep[o.Z] = False
    assert k == pytest.approx(-0.5 * float(np.sum(np.abs(bz[keep]) ** 2)) / o.N, rel=1e-12)
    assert k > e0


def test_initialize_mu_is_a_tenth_scaled_gp_draw():
    """_initialize_mu (MultiResParameters.h:743-759): mu0 = B(F(z) sqrt(lambda)
    / (10 N)) with Z zeroed.  Its prior energy is then -1/2 sum_{k not in Z}
    |F(z)_k|^2 / (100 N) (B B = N x reflection and lambda is even), i.e.
    -1/2 sum |z|^2 / 100 when Z is empty: a GP draw scaled by 1/10."""
    o, _ = _tiny(False, seed=13)
    rng = np.random.default_rng(14)
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)])
    o._initialize_mu()
    assert o.log_prior() == pytest.approx(-0.5 * float(np.sum(zr ** 2 + zi ** 2)) / 100.0, rel=1e-12)
    # with k = 5 the scale is 1/25; with a non-empty Z the energy on Z is dropped
    o.Z = np.array([2, _neg(o, 2)])
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)])
    o._initialize_mu(k=5)
    fz = mrp.F((zr + 1j * zi).reshape(o.shape)).reshape(-1)
    keep = np.ones(o.N, bool)
    keep[o.Z] = False
    lp = -0.5 * float(np.sum(np.abs(fz[keep]) ** 2)) / (25.0 * o.N)
    assert o.log_prior() == pytest.approx(lp, rel=1e-12)
    assert np.allclose(o.mu.reshape(-1)[o.I].real, o.real_mu)


def test_sigma_gibbs_shape_and_rate():
    """_update_sigma_sq_inv (MultiResParameters.h:820-859): Gamma(n_h/2 + 1,
    scale 1 / (||Yh - Re mu[I]||^2 / 2)) then Gamma(n_s/2 + 1, scale
    1 / (||Ys - W Re mu[I]||^2 / 2)); the rates are recomputed here with loops
    over the CSR rows.  _initialize_sigma (:764-778): Gamma(tau^2, 1), then
    sigma_s = sigma_h / U."""
    o, data = _tiny(True, seed=15)
    calls = []

    class Capture:
        def gamma(self, shape, scale):
            calls.append((shape, scale))
            return 1.0 + len(calls)

        def uniform(self):
            return 0.25

    o.rng = Capture()
    mu_re = [o.mu.reshape(-1)[i].real for i in o.I]
    o.real_mu = o._gather_re(o.mu)
    o._update_sigma_sq_inv(data)
    ssh = sum((y - m) ** 2 for y, m in zip(data.Yh.tolist(), mu_re))
    W = data.W
    sss = 0.0
    for r in range(W.shape[0]):
        acc = 0.0
        for t in range(W.indptr[r], W.indptr[r + 1]):
            acc += W.data[t] * mu_re[W.indices[t]]
        sss += (data.Ys[r] - acc) ** 2
    assert calls[0][0] == 0.5 * len(o.I) + 1 and calls[1][0] == 0.5 * W.shape[0] + 1
    assert calls[0][1] == pytest.approx(1.0 / (0.5 * ssh), rel=1e-12)
    assert calls[1][1] == pytest.approx(1.0 / (0.5 * sss), rel=1e-12)
    assert o.sigma_sq_inv == [2.0, 3.0]
    calls.clear()
    o._initialize_sigma(THETA[0])
    assert calls == [(THETA[0], 1.0)]
    assert o.sigma_sq_inv == [2.0, 2.0 / 0.25]


def test_tau_update_matches_rescaled_covariance():
    """The reference's tau^-2 assignment (MultiResParameters.h:870-871)
    multiplies lambda by new / old tau^-2.  The spectrum is linear in the
    variance parameter, so the updated chain equals a fresh chain built with
    variance theta0 * (t / tau0^-2): same eigenvalues, gradient and prior.  The
    Gibbs draw uses shape n/2 + 1 and rate |Re mu at I|^2 / 2 (:866-867)."""
    o, data = _tiny(dual=True)
    t = 3.7
    f = t / o.tau_sq_inv
    o.set_tau_sq_inv(t)
    th = (THETA[0] * f, THETA[1], THETA[2])
    dims = (4, 3, 3)
    ijk = np.array([[i, j, k] for k in range(3) for j in range(3) for i in range(4) if (i + j + k) % 3 != 1])
    ref = mrp.MultiResParameters(2, th, ijk, problem.qform_axis(1.0, dims), rng=mrp.NumpyRng(0))
    ref.sigma_sq_inv = list(o.sigma_sq_inv)
    ref.mu = o.mu.copy()
    np.testing.assert_allclose(o.lam, ref.lam, rtol=1e-12, atol=1e-12 * np.abs(ref.lam).max())
    g, g_ref = o.compute_gradient(data, o.mu), ref.compute_gradient(data, ref.mu)
    np.testing.assert_allclose(g, g_ref, rtol=1e-10, atol=1e-12 * np.abs(g_ref).max())
    assert o.log_prior() == pytest.approx(ref.log_prior(), rel=1e-12)
    assert o.tau() == pytest.approx(1 / np.sqrt(t), rel=1e-15)

    class Capture(mrp.NumpyRng):
        def gamma(self, shape, scale):
            self.args = (shape, scale)
            return 2.5

    o.rng = Capture(1)
    o.real_mu = o._gather_re(o.mu)
    o._update_tau_sq_inv()
    re_at_I = o.mu.reshape(-1)[o.I].real
    assert o.rng.args[0] == pytest.approx(0.5 * len(o.I) + 1)
    assert o.rng.args[1] == pytest.approx(2.0 / np.sum(re_at_I ** 2), rel=1e-12)
    assert o.tau_sq_inv == 2.5
