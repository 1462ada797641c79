"""Host-side sampler driver over a device chain.

Mirrors the reference's scalar control around the hot path:
  * HMCParameters: dual-averaging step size (include/dualres/HMCParameters.h:97-297)
  * mcmc_summary: running moments (include/dualres/gaussian_process_model.h:191-317)
  * fit: the fit_model loop (include/dualres/gaussian_process_model.h:87-177)
The host draws the scalar random numbers the reference draws from its global
std::mt19937 (eps jitter, accept uniform, Gibbs gammas); the bulk normals come
from the device Philox stream.
"""
import math
import time

import numpy as np


class HMCParameters:
    """Dual averaging of Hoffman & Gelman (2014) as modified by the reference."""

    def __init__(self, burnin=1000, n_save=500, thin=1, integrator_steps=10, starting_eps=0.1,
                 mh_target=0.65, gamma=0.05, t0=10.0, kappa=0.75):
        if burnin < 0:
            raise ValueError("HMCParameters: burnin must be >= 0")
        if n_save < 0:
            raise ValueError("HMCParameters: n_save must be >= 0")
        if thin < 0:
            raise ValueError("HMCParameters: thin must be >= 0")
        if integrator_steps <= 0:
            raise ValueError("HMCParameters: integrator_steps must be > 0")
        if starting_eps <= 0:
            raise ValueError("HMCParameters: starting_eps must be > 0")
        if not 0 < mh_target < 1:
            raise ValueError("HMCParameters: mh_target should be within (0, 1)")
        if not 0 < gamma < 1:
            raise ValueError("HMCParameters: gamma should be within (0, 1)")
        if t0 < 0:
            raise ValueError("HMCParameters: t0 should be >= 0")
        if not 0 < kappa < 1:
            raise ValueError("HMCParameters: kappa should be within (0, 1)")
        self._warmup = burnin > 0
        self._eps_start_found = not self._warmup
        self.burnin = burnin
        self._L = integrator_steps
        self.iteration = 0
        self.n_save = n_save
        self.thin = 1 if thin == 0 else thin
        self._A = 1.0
        self._eps0 = 1e-5
        self._eps = max(float(starting_eps), self._eps0)
        self._eps_bar = 1.0
        self._gamma = gamma
        self._h_bar = 0.0
        self._mh_sum = 0.0
        self._mh_target = mh_target
        self._t0 = t0
        self._kappa = kappa
        self._eps_target = 0.0

    def save_iteration(self):
        return self.iteration > self.burnin and (self.iteration - self.burnin) % self.thin == 0

    def integrator_steps(self):
        return self._L if self._eps_start_found else 1

    def max_iterations(self):
        return self.burnin + self.thin * self.n_save

    def eps(self, u09_11):
        if self._warmup:
            return self._eps
        return max(self._eps * u09_11, self._eps0)

    def eps_value(self):
        return self._eps

    def metropolis_hastings_rate(self):
        return self._mh_sum / max(self.iteration - self.burnin, 1)

    def update(self, mh_rate):
        if self.iteration == self.burnin:
            self._warmup = False
            self._eps_start_found = True
            if self._eps_bar > self._eps0:
                self._eps = self._eps_bar
        if self._warmup:
            if not self._eps_start_found:
                if self.iteration == 0:
                    self._A = 2 * int(mh_rate > 0.5) - 1
                if math.pow(mh_rate, self._A) > math.pow(2.0, -self._A):
                    self._eps *= math.pow(2.0, self._A)
                else:
                    self._eps_start_found = True
                    self._eps_target = math.log(10.0 * self._eps)
            else:
                it = self.iteration
                self._h_bar = (1 - 1 / (it + self._t0)) * self._h_bar + (self._mh_target - mh_rate) / (it + self._t0)
                self._eps = max(math.exp(self._eps_target - math.sqrt(it) * self._h_bar / self._gamma), self._eps0)
                a = math.pow(it, -self._kappa)
                self._eps_bar = math.exp(a * math.log(self._eps) + (1 - a) * math.log(self._eps_bar))
        else:
            self._mh_sum += mh_rate
        self.iteration += 1


class McmcSummary:
    """Running posterior moments (gaussian_process_model.h:191-302), float64.

    With ``chain`` given the moments of mu are accumulated on the device
    (``Chain.summary_update``) and ``update`` takes ``mu=None``; otherwise
    they are accumulated here from the host vector ``mu``."""

    def __init__(self, n_mu, n_sigma, chain=None):
        self.samples = 0
        self.chain = chain
        self._m = None
        if chain is not None:
            chain.summary_reset()
        else:
            self._m = (np.zeros(n_mu), np.zeros(n_mu), np.zeros(n_mu))
        self.sigma = np.zeros(n_sigma)
        self.log_lik = 0.0

    def update(self, mu, sigma, log_lik):
        if self.chain is not None:
            self.chain.summary_update()
            self._m = None
        else:
            m1, m2, ma = self._m
            m1 += mu
            ma += np.abs(mu) / max(np.abs(mu).max(), 1e-6)
            m2 += mu * mu
        self.sigma += sigma
        self.log_lik += log_lik
        self.samples += 1

    def _moments(self):
        if self._m is None:
            s1, s2, sa, _ = self.chain.summary_read()
            self._m = (s1, s2, sa)
        return self._m

    @property
    def m1(self):
        return self._moments()[0]

    @property
    def m2(self):
        return self._moments()[1]

    @property
    def m1_abs(self):
        return self._moments()[2]

    def _n(self):
        return max(self.samples, 1)

    def mode_mu(self):
        return self.m1 / self._n()

    def var_mu(self):
        return (self.m2 - self.m1 * self.m1 / self._n()) / self._n()

    def activation(self):
        m = self.m1_abs / (self._n() * np.sqrt(self.var_mu()))
        return m / m.max()


def gibbs_sigma(chain, rng, n_h, n_s, sse_h, sse_s):
    """_update_sigma_sq_inv (MultiResParameters.h:820-859)."""
    s = [rng.gamma(0.5 * n_h + 1, 1.0 / (0.5 * sse_h))]
    if chain.n_datasets == 2:
        s.append(rng.gamma(0.5 * n_s + 1, 1.0 / (0.5 * sse_s)))
    chain.set_sigma_sq_inv(s)
    return s


def gibbs_tau(chain, rng, n_h):
    """_update_tau_sq_inv (MultiResParameters.h:863-872): Gamma(n_h/2 + 1,
    rate = sum mu^2 / 2) over the data voxels; lambda *= new / old tau^-2."""
    t = rng.gamma(0.5 * n_h + 1, 1.0 / (0.5 * chain.mu_data_sumsq()))
    chain.set_tau_sq_inv(t)
    return t


def init_sigma(chain, rng, tau2):
    """_initialize_sigma (MultiResParameters.h:764-778)."""
    s = [rng.gamma(tau2, 1.0)]
    if chain.n_datasets == 2:
        s.append(s[0] / rng.uniform())
    chain.set_sigma_sq_inv(s)
    return s


def fit(chain, hmc, tau2, n_s=0, seed=1, on_iteration=None, update_tau=False):
    """fit_model loop (gaussian_process_model.h:134-171) on a prepared chain.

    The chain must have its data, mu (with data added) and output indices
    set.  Returns (McmcSummary, stats dict).
    """
    rng = np.random.default_rng(seed)
    init_sigma(chain, rng, tau2)
    summ = McmcSummary(chain.n_out, chain.n_datasets, chain=chain)
    save_count = 0
    t_start = time.perf_counter()
    grads = 0
    while save_count < hmc.n_save and hmc.iteration <= hmc.max_iterations():
        eps = hmc.eps(rng.uniform(0.9, 1.1))
        L = hmc.integrator_steps()
        u = rng.uniform()
        r = chain.update(eps, L, u)
        grads += L + 1
        hmc.update(r["mh_rate"])
        s = gibbs_sigma(chain, rng, chain.n_h, n_s, r["sse_h"], r["sse_s"])
        if update_tau:
            gibbs_tau(chain, rng, chain.n_h)
        if hmc.save_iteration():
            if save_count == 0:
                t_start = time.perf_counter()
                grads = 0
            ll = -0.5 * (s[0] * r["sse_h"] + (s[1] * r["sse_s"] if chain.n_datasets == 2 else 0.0))
            summ.update(None, np.sqrt(1.0 / np.asarray(s)), ll)
            save_count += 1
        if on_iteration:
            on_iteration(hmc, r)
    t = time.perf_counter() - t_start
    return summ, {"sampling_time": t, "mh_rate": hmc.metropolis_hastings_rate(), "eps": hmc.eps_value(),
                  "post_burnin_grad_evals": grads}


def fit_native(chain, n_s, tau2, burnin, nsave, thin, L, seed, post_iters=0, captured=False):
    """The C++ host's fit loop (dualgpm's fit_model iterations, in
    libdualres_host.so) on a prepared chain: the sigma initialisation, HMC
    updates, Gibbs draws, dual averaging and device summaries, with the host
    scalars in the chain's precision as dualgpm (float64) / dualgpmf
    (float32) keep them; captured replays each iteration from a captured
    graph.  Returns post-burn-in timing and sampler stats."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdualres_host.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.drh_fit_chain.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    lib.drh_fit_chain.restype = C.c_int
    out = (C.c_double * 7)()
    rc = lib.drh_fit_chain(chain._h, 0 if chain.precision == "fp64" else 1, chain.n_datasets, chain.n_h, int(n_s),
                           chain.n_out, float(tau2), int(burnin), int(nsave), int(thin), int(L), int(seed),
                           int(post_iters), int(bool(captured)), out)
    if rc != 0:
        raise RuntimeError("drh_fit_chain failed")
    return {"wall_s": out[0], "grad_evals": int(out[1]), "mh_rate": out[2], "eps": out[3], "p50_s": out[4],
            "p99_s": out[5], "iterations": int(out[6])}
