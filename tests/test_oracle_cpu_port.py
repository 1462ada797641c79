"""Pins the C++ CPU restatement of the sampler (oracle/cpu_hmc.cpp, the CPU
baseline) against the numpy oracle on the same inputs: its own 3D DFT (radix
4, 2, 3, 5 and 7 stages) against the DFT definition through scipy, the
spectrum's nonpositive count, the gradient, a leapfrog trajectory with its
energies and a full update with injected normals, in float64 (1e-10) and
float32 (tolerances of single-precision rounding), single and dual
resolution, with single-threaded and OpenMP-split W products."""
import numpy as np
import pytest

from helpers import oracle_chain, random_state, rel, small_dual, small_single
from oracle import cpu_port, mrp

THETA1 = (0.806, 0.132, 1.0)
THETA2 = (0.806, 0.132, 1.5)


def _port(p, theta, o, precision, threads=2, w_threads=1):
    return cpu_port.CpuChain(p["ijk"], theta, p["q_hr"], o.M, p["Yh"], p.get("Ys"), p.get("W"),
                             precision=precision, threads=threads, w_threads=w_threads, seed=5)


@pytest.mark.parametrize("embed", [(24, 20, 18), (28, 30, 14)])
def test_cpu_dft_matches_definition(embed):
    """Gradient of a state with no data term reduces to -F(B(mu)/(N lambda)),
    so the port's DFT pair is checked through the oracle's scipy transforms on
    lengths 24 = 4 2 3, 20 = 4 5, 18 = 2 3 3, 28 = 4 7, 30 = 2 3 5, 14 = 2 7."""
    p = small_single(dims=(8, 7, 6))
    o, data = oracle_chain(p, THETA1, embed=embed)
    c = _port(p, THETA1, o, "fp64", threads=3)
    assert c.nonpositive() == o.Z.size
    rng = np.random.default_rng(0)
    mu, mom = random_state(o, rng)
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(0.0)]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv([0.0])
    assert rel(c.gradient(), o.compute_gradient(data, o.mu)) < 1e-12


@pytest.mark.parametrize("dual,w_threads", [(False, 1), (True, 1), (True, 3)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cpu_port_matches_oracle(dual, w_threads, precision):
    p = small_dual() if dual else small_single()
    theta = THETA2 if dual else THETA1
    o, data = oracle_chain(p, theta)
    c = _port(p, theta, o, precision, w_threads=w_threads)
    rng = np.random.default_rng(3)
    mu, mom = random_state(o, rng)
    mu.reshape(-1)[o.I] += p["Yh"]
    sig = [1.3, 0.7][: o.n_datasets]
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(s) for s in sig]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv(sig)
    tol = 1e-10 if precision == "fp64" else 1e-4
    assert rel(c.gradient(), o.compute_gradient(data, o.mu)) < tol
    o._compute_mass_matrix_eigen_values()
    lp0, k0 = o._log_posterior(data, mu), o._potential_energy(mom)
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), 0.07, 6)
    e = c.leapfrog(0.07, 6)
    m_c, p_c = c.get_state()
    ttol = 1e-9 if precision == "fp64" else 1e-3
    assert rel(m_c, mu1) < ttol and rel(p_c, p1) < ttol
    etol = 1e-9 if precision == "fp64" else 1e-5
    for got, want in zip(e, (lp0, k0, o._log_posterior(data, mu1), o._potential_energy(p1))):
        assert abs(got - want) < etol * abs(want)
    # full iteration with injected normals and a scripted accept uniform
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    o.mu = m_c.copy()
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)], uniforms=[0.3], gammas=[1.0, 1.0])
    o.update(data, 0.05, 5)
    r = c.update(0.05, 5, 0.3, noise=(zr, zi))
    assert abs(r[0] - o.initial_energy) < etol * abs(o.initial_energy)
    assert abs(r[1] - o.proposed_energy) < etol * abs(o.proposed_energy)
    assert bool(r[3]) == bool(o.accepted)
    c.close()
