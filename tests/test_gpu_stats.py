"""Statistical parity of the device sampler's own random path.

The product path draws its momentum and initial state from a counter-based
Philox4x32-10 stream with Box-Muller normals on the device (the reference
draws them from std::mt19937, MultiResParameters.h:743-759, 783-815, so the
values cannot match draw for draw); the momentum's are drawn directly as the
Hermitian white-noise spectrum F(z) of a white field z.  These tests check
what the reference's algorithm fixes regardless of the generator:

* E0, which the engine sums from the fresh momentum as the end point's
  kinetic energy is summed, equals -1/2 sum |z|^2 of the normals recovered
  from that momentum (Parseval), and those normals are iid N(0, 1),
  independent between the real and imaginary streams and between iterations;
* the initial state is a GP draw scaled by 1/10 (its prior energy);
* with the noise precision held fixed, the chain's marginal means and
  variances at the data voxels match the exact Gaussian posterior within
  Monte-Carlo error (float64 and float32 gradients).
"""
import numpy as np
import pytest
from scipy import stats

from dualres_amd import problem
from helpers import dense_posterior, engine_chain, mc_compare, oracle_chain
from oracle import mrp

pytestmark = pytest.mark.gpu

THETA = (0.806, 0.5, 1.5)  # every eigenvalue of the embeddings below positive


def _recover_normals(c, o, p):
    """z = F(B(p) / sqrt(lambda_m)) / N inverts p = F(sqrt(lambda_m) B(z) / N)."""
    o._compute_mass_matrix_eigen_values()
    return mrp.F(mrp.B(p) / np.sqrt(o.lam_mass)) / o.N


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_philox_momentum_is_standard_normal_and_kinetic_equals_e0(gpu_available, precision):
    """dr_update with eps = 0 leaves the state in place, so the end momentum is
    the fresh draw: its kinetic energy k1 must equal e0 and the normals
    recovered from it must be N(0, 1) (moments, Kolmogorov-Smirnov),
    uncorrelated between Re and Im and across two iterations."""
    p = problem.single_problem((14, 12, 10), 2.0, THETA, seed_gp=5, seed_noise=6)
    embed = (48, 40, 32)
    o, _ = oracle_chain(p, THETA, embed=embed)
    assert o.Z.size == 0
    c = engine_chain(p, THETA, precision=precision, state_mode="compat", embed=embed, seed=77)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.4])
    o.sigma_sq_inv = [np.float64(1.4)]
    N = c.N
    zs = []
    for it in range(2):
        r = c.update(0.0, 1, 0.5)
        assert r["k1"] == pytest.approx(r["e0"], rel=1e-12 if precision == "fp64" else 2e-6)
        # E0 / N is -1/2 E|z|^2 = -1 per torus point, sd 1/sqrt(N)
        assert abs(r["e0"] / N + 1.0) < 5.0 / np.sqrt(N)
        _, pm = c.get_state()
        z = _recover_normals(c, o, pm).reshape(-1)
        zs.append(z)
        assert -0.5 * float(np.sum(np.abs(z) ** 2)) == pytest.approx(r["e0"], rel=1e-9 if precision == "fp64" else 1e-5)
    for z in zs:
        for part in (z.real, z.imag):
            assert abs(part.mean()) < 5.0 / np.sqrt(N)
            assert abs(part.var() - 1.0) < 5.0 * np.sqrt(2.0 / N)
            assert abs(stats.skew(part)) < 5.0 * np.sqrt(6.0 / N)
            assert abs(stats.kurtosis(part)) < 5.0 * np.sqrt(24.0 / N)
            assert stats.kstest(part, "norm").pvalue > 1e-4
        assert abs(np.corrcoef(z.real, z.imag)[0, 1]) < 5.0 / np.sqrt(N)
    for a, b in ((zs[0].real, zs[1].real), (zs[0].imag, zs[1].imag), (zs[0].real, zs[1].imag)):
        assert abs(np.corrcoef(a, b)[0, 1]) < 5.0 / np.sqrt(N)
    # neighbouring torus points are independent too (lag-1 along x, y and z)
    zr = zs[0].real.reshape(c.shape)
    for ax in range(3):
        assert abs(np.corrcoef(zr.ravel(), np.roll(zr, 1, axis=ax).ravel())[0, 1]) < 5.0 / np.sqrt(N)
    c.close()


def test_initial_state_is_a_tenth_scaled_gp_draw(gpu_available):
    """mu0 = B(F(z) sqrt(lambda) / (10 N)) (MultiResParameters.h:743-759): its
    prior energy is -1/2 sum |F(z)|^2 / (100 N), mean -N/100 and sd
    sqrt(N)/100 over the device's complex normals (Z empty)."""
    p = problem.single_problem((20, 16, 12), 2.0, THETA, seed_gp=7, seed_noise=8)
    embed = (48, 40, 32)
    for seed in (1, 2):
        c = engine_chain(p, THETA, precision="fp64", state_mode="compat", embed=embed, seed=seed)
        assert c.n_nonpositive == 0
        N = c.N
        lp = c.log_prior()
        assert abs(lp + N / 100.0) < 5.0 * np.sqrt(N) / 100.0, (lp, -N / 100.0)
        c.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_chain_samples_the_gaussian_posterior(gpu_available, precision):
    """3,000 post-burn-in iterations of dr_update (device Philox momentum, host
    accept uniforms) with sigma held fixed on a 12x10x8 torus: marginal means
    of Re(mu) at the 48 data voxels within Monte-Carlo error of the dense
    posterior mean (max |z| < 4.5, mean z^2 in (0.4, 1.8)), variances within
    5 standard errors; the data-free imaginary chain keeps the prior
    variance tau^2."""
    p = problem.single_problem((6, 5, 4), 2.0, THETA, seed_gp=3, seed_noise=4)
    embed = (12, 10, 8)
    o, _ = oracle_chain(p, THETA, embed=embed)
    m, Sig, C = dense_posterior(o, p["Yh"], 1.0)
    c = engine_chain(p, THETA, precision=precision, state_mode="compat", embed=embed, seed=2027)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.0])
    rng = np.random.default_rng(99)
    re, im = [], []
    acc = 0
    n_burn, n_keep = 300, 3000
    for it in range(n_burn + n_keep):
        r = c.update(0.12, 12, rng.uniform())
        acc += r["accepted"]
        if it >= n_burn:
            mu, _ = c.get_state()
            flat = mu.reshape(-1)[o.I]
            re.append(flat.real.copy())
            im.append(flat.imag.copy())
    assert acc / (n_burn + n_keep) > 0.8
    z, ess, ratio, vz = mc_compare(re, m[o.I], np.diag(Sig)[o.I])
    print(f"{precision}: mean z^2 {np.mean(z ** 2):.2f} max|z| {np.abs(z).max():.2f} min ESS {ess.min():.0f} "
          f"var ratio [{ratio.min():.3f}, {ratio.max():.3f}]", flush=True)
    assert ess.min() > 300
    assert np.abs(z).max() < 4.5 and 0.4 < np.mean(z ** 2) < 1.8
    assert np.abs(vz).max() < 5.0
    z, ess, ratio, vz = mc_compare(im, np.zeros(len(o.I)), np.diag(C)[o.I])
    assert np.abs(z).max() < 4.5 and np.abs(vz).max() < 5.0
    c.close()
