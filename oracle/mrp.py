"""``MultiResParameters<T>`` restated on the CPU (oracle; test only).

Follows include/dualres/MultiResParameters.h line by line: complex state on
the embedding torus, FFTW conventions (F = forward, B = backward, both
unnormalised, x fastest), c2c FFTs everywhere, the low-rank set Z of
nonpositive eigenvalues zeroed after every spectral division.

Random draws are injected through an ``rng`` object with
``complex_normals(n) -> (re, im)``, ``uniform() -> float`` and
``gamma(shape, scale) -> float``, so that a test can script the exact values
the reference would have drawn from its global ``std::mt19937``.

Pinned by (tests/test_oracle_mrp.py): a dense circulant restatement on tiny
tori (g = -C^+ mu + P^T r), central finite differences of log_posterior,
leapfrog time reversibility and O(eps^2) energy error, and the closed-form
harmonic-oscillator map when every torus point is observed.
"""
import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp

from . import spectrum


def F(x):
    """FFTW_FORWARD, unnormalised (sized_fft.h:139-159)."""
    return sfft.fftn(x, axes=(0, 1, 2))


def B(x):
    """FFTW_BACKWARD, unnormalised: B = N * F^{-1} (sized_fft.h:144-148)."""
    return sfft.ifftn(x, axes=(0, 1, 2), norm="forward")


def data_indices(ijk, M):
    """__lambda_grid_indices  (MultiResParameters.h:287-289)."""
    ijk = np.asarray(ijk, dtype=np.int64)
    return ijk[:, 2] * M[0] * M[1] + ijk[:, 1] * M[0] + ijk[:, 0]


class MultiResData:
    """Data container: Yh, Ys, W (row-major CSR)  (MultiResData.h:21-127)."""

    def __init__(self, Yh, Ys=None, W=None, dtype=np.float64):
        self.Yh = np.asarray(Yh, dtype=dtype)
        if Ys is None:
            self.n_datasets = 1
            self.Ys = np.zeros(1, dtype=dtype)
            self.W = None
        else:
            self.n_datasets = 2
            self.Ys = np.asarray(Ys, dtype=dtype)
            self.W = sp.csr_matrix(W, dtype=dtype)


class MultiResParameters:
    """Restatement of dualres::MultiResParameters<T> (MultiResParameters.h:32-157)."""

    def __init__(self, n_datasets, theta, ijk, qform, rng, covfun="rbf",
                 dtype=np.float64, M=None):
        # MultiResParameters.h:172-300
        self.T = np.dtype(dtype).type
        self.C = np.complex128 if self.T == np.float64 else np.complex64
        self.n_datasets = n_datasets
        ijk = np.asarray(ijk, dtype=np.int64)
        max_ijk = ijk.max(axis=0)
        self.M = tuple(M) if M is not None else spectrum.embedding_dims_pow2(max_ijk)
        for a in range(3):
            if self.M[a] < 2 * max_ijk[a] or self.M[a] % 2:
                raise ValueError("embedding must be even and >= 2*max_ijk")
        self.N = int(np.prod(self.M))
        self.shape = (self.M[2], self.M[1], self.M[0])
        self.lam = spectrum.eigenvalues(self.M, qform, theta, covfun, self.T).astype(self.C)
        self.tau_sq_inv = self.T(1) / self.T(theta[0])
        self.Z = spectrum.check_nonpositive_count(self.lam)
        self.I = data_indices(ijk, self.M)
        self.out_idx = self.I.copy()
        self.rng = rng
        self._initialize_mu()
        self._initialize_sigma(theta[0])

    # -- helpers ---------------------------------------------------------
    def _low_rank_adjust(self, A):
        """Zero Z entries  (MultiResParameters.h:309-319)."""
        A.reshape(-1)[self.Z] = 0
        return A

    def _gather_re(self, mu):
        return mu.reshape(-1)[self.I].real.astype(self.T)

    # -- initialisation ---------------------------------------------------
    def _initialize_mu(self, k=10):
        """MultiResParameters.h:743-759."""
        re, im = self.rng.complex_normals(self.N)
        mu = (np.asarray(re, self.T) + 1j * np.asarray(im, self.T)).astype(self.C)
        mu = mu.reshape(self.shape)
        mu = F(mu).astype(self.C)
        mu = (mu * (np.sqrt(self.lam) / (self.T(k) * self.T(self.N)))).astype(self.C)
        self._low_rank_adjust(mu)
        self.mu = B(mu).astype(self.C)
        self.real_mu = self._gather_re(self.mu)

    def _initialize_sigma(self, shape, rate=1.0):
        """MultiResParameters.h:764-778."""
        if self.n_datasets <= 0:
            raise ValueError("MultiResParameters: n_datasets not set")
        s = [self.T(self.rng.gamma(shape, 1.0 / rate))]
        for _ in range(1, self.n_datasets):
            s.append(self.T(s[0] / self.T(self.rng.uniform())))
        self.sigma_sq_inv = s

    def add_data_to_mu(self, data):
        """MultiResParameters.h:631-641."""
        flat = self.mu.reshape(-1)
        flat[self.I] += data.Yh.astype(self.T)
        self.real_mu = self.real_mu + data.Yh.astype(self.T)

    def set_output_indices(self, ijk):
        """MultiResParameters.h:646-667."""
        ijk = np.asarray(ijk, dtype=np.int64)
        if ijk.ndim != 2 or ijk.shape[1] != 3:
            raise ValueError("MultiResParameters:set_output_indices NCOL")
        if (ijk.max(axis=0) > np.asarray(self.M)).any():
            raise ValueError("MultiResParameters:set_output_indices MAX")
        if (ijk.min(axis=0) < 0).any():
            raise ValueError("MultiResParameters:set_output_indices MIN")
        self.out_idx = data_indices(ijk, self.M)

    def mu_(self):
        """Re(mu) at output indices  (MultiResParameters.h:402-408)."""
        return self.mu.reshape(-1)[self.out_idx].real.astype(self.T)

    def sigma(self, which):
        return np.sqrt(self.T(1) / self.sigma_sq_inv[which])

    def tau(self):
        return np.sqrt(self.T(1) / self.tau_sq_inv)

    # -- gradient ---------------------------------------------------------
    def _likelihood_residual(self, data, real_mu):
        """_real_sub_grad  (MultiResParameters.h:357-368)."""
        s = self.sigma_sq_inv
        r = s[0] * (data.Yh - real_mu)
        if self.n_datasets == 2:
            v = data.Ys - data.W @ real_mu
            r = r + s[1] * (v @ data.W)
        elif self.n_datasets != 1:
            raise ValueError("Gradient only implemented for single or dual resolution")
        return r.astype(self.T)

    def compute_gradient(self, data, mu_star):
        """_compute_gradient  (MultiResParameters.h:337-376)."""
        t = B(mu_star).astype(self.C)
        t = (t / (-self.T(self.N) * self.lam)).astype(self.C)
        self._low_rank_adjust(t)
        t = F(t).astype(self.C)
        self.real_mu = self._gather_re(mu_star)
        r = self._likelihood_residual(data, self.real_mu)
        flat = t.reshape(-1)
        flat[self.I] += r
        return t

    def gradient(self, data):
        return self.compute_gradient(data, self.mu)

    # -- energies ---------------------------------------------------------
    def _log_prior(self, mu_star):
        """MultiResParameters.h:428-476 (sum over Re(lambda) > 0 of |B mu|^2/lambda)."""
        t = B(mu_star).astype(self.C).reshape(-1)
        lam = self.lam.reshape(-1)
        keep = lam.real > 0
        terms = (np.conj(t[keep]).astype(np.complex128) * t[keep].astype(np.complex128)
                 / lam[keep].astype(np.complex128))
        lp = terms.sum()
        return -0.5 * (lp.real + lp.imag) / self.N

    def _log_likelihood(self, data, mu_star):
        """MultiResParameters.h:488-514."""
        self.real_mu = self._gather_re(mu_star)
        d = (data.Yh - self.real_mu).astype(np.float64)
        ll = float(self.sigma_sq_inv[0]) * float(d @ d)
        if self.n_datasets == 2:
            e = (data.Ys - data.W @ self.real_mu).astype(np.float64)
            ll += float(self.sigma_sq_inv[1]) * float(e @ e)
        elif self.n_datasets > 2:
            raise ValueError("Likelihood only implemented for single or dual resolution")
        return -0.5 * ll

    def _log_posterior(self, data, mu_star):
        return self._log_prior(mu_star) + self._log_likelihood(data, mu_star)

    def log_prior(self):
        return self._log_prior(self.mu)

    def log_likelihood(self, data):
        return self._log_likelihood(data, self.mu)

    def log_posterior(self, data):
        return self._log_posterior(data, self.mu)

    def _potential_energy(self, momentum):
        """Kinetic term, called _potential_energy in the reference (:551-603)."""
        t = B(momentum).astype(self.C).reshape(-1)
        lm = self.lam_mass.reshape(-1)
        keep = lm != 0
        terms = (np.conj(t[keep]).astype(np.complex128) * t[keep].astype(np.complex128)
                 / lm[keep].astype(np.complex128))
        pe = terms.sum()
        return -0.5 * (pe.real + pe.imag) / self.N

    # -- HMC --------------------------------------------------------------
    def _compute_mass_matrix_eigen_values(self):
        """MultiResParameters.h:734-738."""
        self.lam_mass = (self.T(1) / self.lam + self.sigma_sq_inv[0]).astype(self.C)
        self._low_rank_adjust(self.lam_mass)

    def _sample_momentum(self):
        """MultiResParameters.h:783-815; returns E0 = -1/2 sum |z|^2."""
        re, im = self.rng.complex_normals(self.N)
        re = np.asarray(re, self.T)
        im = np.asarray(im, self.T)
        e0 = -0.5 * float(np.sum(re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2))
        p = (re + 1j * im).astype(self.C).reshape(self.shape)
        p = B(p).astype(self.C)
        p = (p * (np.sqrt(self.lam_mass) / self.T(self.N))).astype(self.C)
        self._low_rank_adjust(p)
        self.momentum = F(p).astype(self.C)
        return e0

    def leapfrog(self, data, mu_star, momentum, eps, L):
        """Leapfrog integration as in update() (MultiResParameters.h:690-700)."""
        eps = self.T(eps)
        k = self.T(0.5)
        momentum = (momentum + k * eps * self.compute_gradient(data, mu_star)).astype(self.C)
        for step in range(L):
            k = self.T(0.5) if step == L - 1 else self.T(1)
            t = B(momentum).astype(self.C)
            t = (t / (self.lam_mass * (self.T(self.N) / eps))).astype(self.C)
            self._low_rank_adjust(t)
            t = F(t).astype(self.C)
            mu_star = (mu_star + t).astype(self.C)
            momentum = (momentum + k * eps * self.compute_gradient(data, mu_star)).astype(self.C)
        return mu_star, momentum

    def update(self, data, eps, L, update_tau=False):
        """One HMC iteration  (MultiResParameters.h:673-727).  Returns min(1, R).

        ``update_tau`` compiles in the tau^-2 draw that both reference
        binaries leave out (DUALRES_DONT_UPDATE_TAU, :723-725)."""
        mu_star = self.mu.copy()
        self._compute_mass_matrix_eigen_values()
        e0 = self._sample_momentum()
        lp_initial = self._log_posterior(data, self.mu)
        self.initial_energy = e0 + lp_initial
        mu_star, self.momentum = self.leapfrog(data, mu_star, self.momentum, eps, L)
        lp_proposal = self._log_posterior(data, mu_star)
        self.proposed_energy = lp_proposal + self._potential_energy(self.momentum)
        with np.errstate(over="ignore", invalid="ignore"):
            R = float(np.exp(self.proposed_energy - self.initial_energy))
        if np.isnan(R):
            R = 0.0
        self.accepted = self.rng.uniform() < R
        if self.accepted:
            self.mu = mu_star
        self.real_mu = self._gather_re(self.mu)
        self._update_sigma_sq_inv(data)
        if update_tau:
            self._update_tau_sq_inv()
        return min(1.0, R)

    def _update_tau_sq_inv(self):
        """Gibbs draw of tau^-2  (MultiResParameters.h:863-872), then
        ``_lambda *= tau_sq_inv_star / _tau_sq_inv`` in the reference's direction."""
        shape = self.T(0.5) * self.T(self.real_mu.size) + self.T(1)
        rate = self.T(0.5) * self.T(np.dot(self.real_mu, self.real_mu))
        self.set_tau_sq_inv(self.T(self.rng.gamma(shape, 1.0 / rate)))

    def set_tau_sq_inv(self, t):
        """The assignment half of _update_tau_sq_inv (:870-871)."""
        self.lam = (self.lam * (self.T(t) / self.tau_sq_inv)).astype(self.C)
        self.tau_sq_inv = self.T(t)

    def _update_sigma_sq_inv(self, data):
        """Gibbs draws  (MultiResParameters.h:820-859)."""
        shape = self.T(0.5) * self.T(data.Yh.size) + self.T(1)
        d = (data.Yh - self.real_mu)
        rate = self.T(0.5) * self.T(np.dot(d, d))
        self.sigma_sq_inv[0] = self.T(self.rng.gamma(shape, 1.0 / rate))
        if self.n_datasets == 2:
            shape = self.T(0.5) * self.T(data.Ys.size) + self.T(1)
            e = data.Ys - data.W @ self.real_mu
            rate = self.T(0.5) * self.T(np.dot(e, e))
            self.sigma_sq_inv[1] = self.T(self.rng.gamma(shape, 1.0 / rate))
        elif self.n_datasets != 1:
            raise ValueError("Parameter updates not implemented for > 2 datasets")


class NumpyRng:
    """Adapter over numpy.random.Generator with the draw methods the oracle needs."""

    def __init__(self, seed):
        self.g = np.random.default_rng(seed)

    def complex_normals(self, n):
        return self.g.standard_normal(n), self.g.standard_normal(n)

    def uniform(self):
        return float(self.g.random())

    def gamma(self, shape, scale):
        return float(self.g.gamma(shape, scale))


class ScriptedRng:
    """Returns pre-set draws in order (for deterministic parity tests)."""

    def __init__(self, normals=(), uniforms=(), gammas=()):
        self.normals = list(normals)
        self.uniforms = list(uniforms)
        self.gammas = list(gammas)

    def complex_normals(self, n):
        re, im = self.normals.pop(0)
        assert len(re) == n and len(im) == n
        return re, im

    def uniform(self):
        return self.uniforms.pop(0)

    def gamma(self, shape, scale):
        g = self.gammas.pop(0)
        return g(shape, scale) if callable(g) else g
