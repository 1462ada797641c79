"""Captured HMC iterations and batches of uneven chains, against the eager
update (MultiResParameters.h:673-727 as dr_update runs it).

A captured iteration replays the same kernels with the same scalars, read
from device memory, so every result must be bitwise equal to dr_update's; a
batch must give each chain exactly what it gets alone, in every mode."""
import numpy as np
import pytest

from dualres_amd import Chain, problem
from dualres_amd.batch import Batch
from helpers import engine_chain, small_dual, small_single

pytestmark = pytest.mark.gpu

THETA = (0.806, 0.132, 1.5)


def _chain(p, precision, seed, mode="compat"):
    c = engine_chain(p, THETA, precision=precision, state_mode=mode, seed=seed)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.1, 0.6][: c.n_datasets])
    return c


def _same(a, b):
    for k in ("H0", "H1", "R", "mh_rate", "sse_h", "sse_s", "accepted"):
        assert a[k] == b[k] or (np.isnan(a[k]) and np.isnan(b[k])), (k, a[k], b[k])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dual", [False, True])
def test_captured_update_is_bitwise_the_eager_update(gpu_available, precision, dual):
    """Eight iterations with varying step sizes, leapfrog lengths, noise
    precisions and accept uniforms (half accepted whatever the energy, the
    others rejected unless it fell): energies, accept
    decisions and the kept state equal dr_update's bit for bit."""
    p = small_dual(dims=(16, 14, 12)) if dual else small_single(dims=(14, 12, 10))
    a, b = _chain(p, precision, 9), _chain(p, precision, 9)
    rng = np.random.default_rng(4)
    n_acc = 0
    for it in range(8):
        # u = 0 accepts any finite proposal; u = 0.9999 rejects unless the energy rose
        eps, L, u = rng.uniform(0.02, 0.15), int(rng.integers(1, 7)), 0.0 if it % 2 == 0 else 0.9999
        r1 = a.update(eps, L, u)
        r2 = b.update_captured(eps, L, u)
        _same(r1, r2)
        n_acc += r1["accepted"]
        s = (rng.gamma(10.0, 0.1), rng.gamma(5.0, 0.1))[: a.n_datasets]
        a.set_sigma_sq_inv(list(s))
        b.set_sigma_sq_inv(list(s))
    assert 0 < n_acc < 8
    assert np.array_equal(a.mu(), b.mu())
    ma, pa = a.get_state()
    mb, pb = b.get_state()
    assert np.array_equal(ma, mb)
    a.close()
    b.close()


def _uneven_problems():
    out = []
    for k, dims in enumerate(((14, 12, 10), (20, 16, 12), (16, 22, 14), (24, 18, 20))):
        if k % 2:
            out.append(small_dual(dims=dims, seed=11 + k))
        else:
            out.append(small_single(dims=dims, seed=11 + k))
    return out


@pytest.mark.parametrize("mode", ["graph", "streams", "eager"])
def test_batch_chains_equal_alone(gpu_available, mode):
    """Four chains of uneven size (single and dual resolution, float32 and
    float64) in one batch: each chain's energies, accept decisions and state
    after six iterations equal the same chain updated alone."""
    probs = _uneven_problems()
    precs = ["fp32", "fp64", "fp32", "fp32"]
    alone = [_chain(p, pr, 30 + i) for i, (p, pr) in enumerate(zip(probs, precs))]
    inb = [_chain(p, pr, 30 + i) for i, (p, pr) in enumerate(zip(probs, precs))]
    rng = np.random.default_rng(8)
    with Batch(inb, mode) as bt:
        for it in range(6):
            eps = rng.uniform(0.05, 0.3, size=4)
            L = rng.integers(1, 6, size=4) if it < 3 else np.full(4, 5)
            u = rng.uniform(size=4)
            rb = bt.update(eps, L, u)
            for i in range(4):
                ra = alone[i].update(eps[i], int(L[i]), u[i])
                _same(ra, rb[i])
            for i in range(4):
                s = [rng.gamma(10.0, 0.1), rng.gamma(5.0, 0.1)][: alone[i].n_datasets]
                alone[i].set_sigma_sq_inv(s)
                inb[i].set_sigma_sq_inv(s)
            # the chains' own calls work between batch updates
            for c in inb:
                c.summary_update()
    for a, b in zip(alone, inb):
        assert np.array_equal(a.mu(), b.mu())
        assert b.summary_read()[3] == 6
    for c in alone + inb:
        c.close()


def test_batch_of_one_and_reuse_after_batch(gpu_available):
    """A chain leaves a destroyed batch able to update on its own again, and
    its sequence continues exactly as an unbatched chain's."""
    p = small_dual(dims=(16, 14, 12))
    a, b = _chain(p, "fp32", 5), _chain(p, "fp32", 5)
    for it in range(3):
        ra = a.update(0.1, 4, 0.5)
        if it == 0:
            with Batch([b], "graph") as bt:
                rb = bt.update([0.1], [4], [0.5])[0]
        else:
            rb = b.update(0.1, 4, 0.5)
        _same(ra, rb)
    assert np.array_equal(a.mu(), b.mu())
    a.close()
    b.close()


def test_fit_batch_equals_fit_alone(gpu_available):
    """fit_batch (lockstep fits through one graph per iteration) against fit()
    on each chain alone: same dual-averaged step sizes, posterior means and
    variances, bit for bit, over burn-in (eps search with L = 1, then dual
    averaging) and saved iterations."""
    from dualres_amd.batch import fit_batch
    from dualres_amd.sampler import HMCParameters, fit
    probs = _uneven_problems()[:3]
    seeds = [71, 72, 73]
    n_s = [p["Ys"].size if "Ys" in p else 0 for p in probs]

    def hmc():
        return HMCParameters(burnin=8, n_save=5, thin=1, integrator_steps=5, starting_eps=0.05)

    alone = [_chain(p, "fp32", 40 + i) for i, p in enumerate(probs)]
    inb = [_chain(p, "fp32", 40 + i) for i, p in enumerate(probs)]
    ref = [fit(c, hmc(), THETA[0], n_s=n_s[i], seed=seeds[i]) for i, c in enumerate(alone)]
    hb = [hmc() for _ in probs]
    summ, stats, info = fit_batch(inb, hb, [THETA[0]] * 3, n_s, seeds, mode="graph")
    assert info["post_burnin_grad_evals"] > 0
    for (sa, st), sb, stb in zip(ref, summ, stats):
        assert st["eps"] == stb["eps"]
        assert np.array_equal(sa.mode_mu(), sb.mode_mu())
        assert np.array_equal(sa.var_mu(), sb.var_mu())
    for c in alone + inb:
        c.close()
