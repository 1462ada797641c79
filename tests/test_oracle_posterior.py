"""Statistical pin of the oracle's HMC iteration (MultiResParameters.h:673-727):
with the noise precision held fixed, a single-resolution posterior is
Gaussian, so a long chain's marginal means and variances at the data voxels
must match the exact posterior (dense inverse on a 12x10x8 torus) within
Monte-Carlo error.  The imaginary half of the complex state carries no data,
so its draws must follow the prior: mean 0, variance tau^2."""
import numpy as np

from dualres_amd import problem
from helpers import dense_posterior, mc_compare
from oracle import mrp

THETA = (0.806, 0.5, 1.5)  # every eigenvalue of the 12x10x8 embedding positive


def test_oracle_chain_samples_the_gaussian_posterior():
    p = problem.single_problem((6, 5, 4), 2.0, THETA, seed_gp=3, seed_noise=4)
    o = mrp.MultiResParameters(1, THETA, p["ijk"], p["q_hr"], rng=mrp.NumpyRng(1), M=(12, 10, 8))
    data = mrp.MultiResData(p["Yh"])
    m, Sig, C = dense_posterior(o, p["Yh"], 1.0)
    o.add_data_to_mu(data)
    o.sigma_sq_inv = [1.0]
    o._update_sigma_sq_inv = lambda d: None   # sigma held fixed
    re, im = [], []
    acc = 0
    for it in range(3300):
        o.update(data, 0.12, 12)
        acc += o.accepted
        if it >= 300:
            flat = o.mu.reshape(-1)[o.I]
            re.append(flat.real.copy())
            im.append(flat.imag.copy())
    assert acc / 3300 > 0.8
    z, ess, ratio, vz = mc_compare(re, m[o.I], np.diag(Sig)[o.I])
    assert ess.min() > 300
    assert np.abs(z).max() < 4.5 and 0.4 < np.mean(z ** 2) < 1.8
    assert np.abs(vz).max() < 5.0
    z, ess, ratio, vz = mc_compare(im, np.zeros(len(o.I)), np.diag(C)[o.I])
    assert np.abs(z).max() < 4.5 and np.abs(vz).max() < 5.0
    assert np.allclose(np.diag(C), THETA[0])
