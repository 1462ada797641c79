"""Dual-averaging step-size control (oracle; test only).

Restates include/dualres/HMCParameters.h:97-297 (Hoffman & Gelman 2014,
Alg. 5, as modified by the reference: an initial doubling/halving search with
L = 1, eps floor 1e-5, eps_bar adopted at the end of burn-in, eps jittered by
U(0.9, 1.1) afterwards).  Deterministic given the sequence of MH rates and
jitter draws; pinned by hand-worked steps and invariants in
tests/test_oracle_hmc.py.
"""
import math


class HMCParameters:
    def __init__(self, burnin=1000, n_save=500, thin=1, integrator_steps=10,
                 starting_eps=0.1, mh_target=0.65, gamma=0.05, t0=10, kappa=0.75):
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
        self.warmup = burnin > 0
        self.eps_start_found = not self.warmup
        self.burnin = burnin
        self.L = integrator_steps
        self.iteration = 0
        self.n_save = n_save
        self.thin = 1 if thin == 0 else thin
        self.A = 1.0
        self.eps0 = 1e-5
        self.eps = max(float(starting_eps), self.eps0)
        self.eps_bar = 1.0
        self.gamma = gamma
        self.h_bar = 0.0
        self.mh_running_sum = 0.0
        self.mh_target = mh_target
        self.t0 = t0
        self.kappa = kappa
        self.eps_target = 0.0

    def save_iteration(self):
        return self.iteration > self.burnin and (self.iteration - self.burnin) % self.thin == 0

    def integrator_steps(self):
        return self.L if self.eps_start_found else 1

    def max_iterations(self):
        return self.burnin + self.thin * self.n_save

    def eps_value(self):
        return self.eps

    def eps_draw(self, u09_11):
        """eps() given a U(0.9, 1.1) draw (HMCParameters.h:212-221)."""
        if self.warmup:
            return self.eps
        return max(self.eps * u09_11, self.eps0)

    def metropolis_hastings_rate(self):
        return self.mh_running_sum / max(self.iteration - self.burnin, 1)

    def update(self, mh_rate):
        """HMCParameters.h:259-297."""
        if self.iteration == self.burnin:
            self.warmup = False
            self.eps_start_found = True
            if self.eps_bar > self.eps0:
                self.eps = self.eps_bar
        if self.warmup:
            if not self.eps_start_found:
                if self.iteration == 0:
                    self.A = 2 * int(mh_rate > 0.5) - 1
                if math.pow(mh_rate, self.A) > math.pow(2.0, -self.A):
                    self.eps *= math.pow(2.0, self.A)
                else:
                    self.eps_start_found = True
                    self.eps_target = math.log(10.0 * self.eps)
            else:
                it = self.iteration
                self.h_bar = ((1 - 1 / (it + self.t0)) * self.h_bar
                              + (self.mh_target - mh_rate) / (it + self.t0))
                self.eps = math.exp(self.eps_target - math.sqrt(it) * self.h_bar / self.gamma)
                self.eps = max(self.eps, self.eps0)
                a = math.pow(it, -self.kappa)
                self.eps_bar = math.exp(a * math.log(self.eps) + (1 - a) * math.log(self.eps_bar))
        else:
            self.mh_running_sum += mh_rate
        self.iteration += 1
