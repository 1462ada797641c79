"""The FFT pass autotuner (dr_autotune): tuned chains keep oracle parity, a
tuning applied by hand reproduces the tuned chain bit for bit, and the JSON
cache returns the timed choice for the same configuration."""
import json

import numpy as np
import pytest

from helpers import engine_chain, oracle_chain, random_state, rel, small_dual
from oracle import mrp

pytestmark = pytest.mark.gpu

THETA = (0.806, 0.132, 1.5)
EMBED = (30, 28, 24)  # x pass length 15 = 3 5, y 28 = 4 7, z 24 = 8 3: runtime stages on every axis


def _setup(p, precision, params=None, seed=0):
    o, data = oracle_chain(p, THETA, embed=EMBED)
    c = engine_chain(p, THETA, precision=precision, state_mode="compat", embed=EMBED, seed=3)
    if params is not None:
        c.set_tuning(params)
    rng = np.random.default_rng(seed)
    mu, mom = random_state(o, rng)
    mu.reshape(-1)[o.I] += p["Yh"]
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(1.3), np.float64(0.7)]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv([1.3, 0.7])
    return o, data, c, mu, mom


@pytest.mark.parametrize("params", [None, [0, 0, 0, 0, 0, 1, 2, 1], [1, 2, 1, 0, 2, 2, 1, 2]])
def test_tuned_shapes_keep_parity(gpu_available, params):
    """Float64 gradient (1e-10) and 5-step leapfrog (1e-9) against the oracle
    with the autotuner's choice (None) and two hand-set extremes of tile lines
    and radix orders."""
    p = small_dual(dims=(16, 14, 12))
    o, data, c, mu, mom = _setup(p, "fp64", params)
    if params is None:
        got, hit = c.autotune()
        assert not hit and len(got) == 8 and c.tuning() == got
        c.set_state(mu, mom)
    assert rel(c.gradient().real, o.compute_gradient(data, o.mu).real) < 1e-10
    o._compute_mass_matrix_eigen_values()
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), 0.05, 5)
    c.leapfrog(0.05, 5)
    m, pm = c.get_state()
    assert rel(m.real, mu1.real) < 1e-9 and rel(pm.real, p1.real) < 1e-9 and rel(m.imag, mu1.imag) < 1e-9
    c.close()


def test_tuning_reproduces_bitwise_and_caches(gpu_available, tmp_path):
    """A chain given the tuned parameters by dr_set_tuning updates exactly as
    the tuned chain does; the cache file holds the timed entry under this
    configuration's key and a second chain of the same configuration reads
    it back instead of timing."""
    p = small_dual(dims=(16, 14, 12))
    cache = tmp_path / "tune.json"
    a = engine_chain(p, THETA, precision="fp32", state_mode="compat", embed=EMBED, seed=8)
    prm, hit = a.autotune(cache)
    assert not hit
    entries = json.loads(cache.read_text())
    assert len(entries) == 1
    key, val = next(iter(entries.items()))
    assert val == prm and "fp32|compat|2|30x28x24|" in key
    b = engine_chain(p, THETA, precision="fp32", state_mode="compat", embed=EMBED, seed=8)
    prm_b, hit_b = b.autotune(cache)
    assert hit_b and prm_b == prm
    for c in (a, b):
        c.add_data_to_mu()
        c.set_sigma_sq_inv([1.1, 0.6])
    rng = np.random.default_rng(2)
    for _ in range(4):
        eps, u = rng.uniform(0.05, 0.2), rng.uniform()
        ra, rb = a.update(eps, 4, u), b.update_captured(eps, 4, u)
        for k in ("H0", "H1", "accepted"):
            assert ra[k] == rb[k]
    assert np.array_equal(a.mu(), b.mu())
    a.close()
    b.close()
