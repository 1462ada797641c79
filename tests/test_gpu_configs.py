"""GPU parity at the benchmark workloads' full sizes, against the CPU oracle.

* Single-resolution 2 mm fit: 40x48x40 image, ellipsoid mask (29,400 data
  voxels), tau2=0.806 psi=0.132 nu=1, float64, on the 80x96x80 embedding and on
  the reference's power-of-two rule (128^3): gradient, an L=25 leapfrog
  trajectory with every energy, and a full update with injected normals.
* Dual-resolution 1 mm fit: highres 128x160x128 (1,000,560 data voxels) and
  stdres 64x80x64 (125,064 rows), neighbourhood 6.9 mm, W built on the device
  (159 M nonzeros, 1,357-entry interior stencil, 880 interior W^T v tiles),
  embedding 256x320x256.  These are the exact kernels the throughput
  benchmark times: x passes of length 128, y passes of length 320, the fused
  z pass of length 256.  Kriging pattern bit-exact against the oracle's
  integer search; gradient and a 3-step leapfrog in float64 and float32;
  stencil layout against plain CSR; a float32 chain against a float64 chain
  on the same draws.
* Embeddings with a factor of 7 and transform lengths 160 and 512, with a
  last z-pass tile that is only partly filled.

Tolerances (relative 2-norm against the float64 oracle): float64 gradient
1e-10, trajectory and energies 1e-9; float32 gradient 2e-4 (1 mm: the W
products sum up to 1,357 float terms per row), trajectory 2e-3, energies
1e-4.  The oracle's FFTs run on up to 16 host threads.
"""
import os

import numpy as np
import pytest
import scipy.fft as sfft

from helpers import engine_chain, oracle_chain, random_state, rel
from oracle import mrp

pytestmark = pytest.mark.gpu

THETA1 = (0.806, 0.132, 1.0)
THETA2 = (0.806, 0.132, 1.5)
WORKERS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(autouse=True)
def _fft_workers():
    with sfft.set_workers(WORKERS):
        yield


def _seeded_state(o, p, c, sig, seed):
    rng = np.random.default_rng(seed)
    mu, mom = random_state(o, rng, imag=(c.state_mode == "compat"))
    mu.reshape(-1)[o.I] += p["Yh"]
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(s) for s in sig[: o.n_datasets]]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv(list(sig[: o.n_datasets]))
    return mu, mom


def _check_leapfrog(o, data, c, mu, mom, eps, L, tol, etol):
    o._compute_mass_matrix_eigen_values()
    lp0 = o._log_posterior(data, mu)
    k0 = o._potential_energy(mom)
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), eps, L)
    lp1 = o._log_posterior(data, mu1)
    k1 = o._potential_energy(p1)
    r = c.leapfrog(eps, L)
    mu_e, p_e = c.get_state()
    errs = (rel(mu_e.real, mu1.real), rel(p_e.real, p1.real), rel(mu_e.imag, mu1.imag))
    print(f"leapfrog L={L}: state {errs[0]:.2e} momentum {errs[1]:.2e} imag {errs[2]:.2e}", flush=True)
    assert max(errs) < tol
    assert abs((r["log_prior1"] + r["log_lik1"]) - lp1) < etol * abs(lp1)
    assert abs((r["log_prior0"] + r["log_lik0"]) - lp0) < etol * abs(lp0)
    assert abs(r["k1"] - k1) < etol * abs(k1)
    assert abs(r["e0"] - k0) < etol * abs(k0)


# ---------------------------------------------------------------- 2 mm single resolution

@pytest.fixture(scope="module")
def cfg1():
    from dualres_amd import problem
    return problem.single_problem((40, 48, 40), 2.0, THETA1, seed_gp=20261202, seed_noise=1)


@pytest.mark.parametrize("embed", [(80, 96, 80), None])
def test_single_res_2mm_full_size_fp64(gpu_available, cfg1, embed):
    """Gradient, L=25 leapfrog with all four energies, and one full update with
    injected normals, at the 2 mm workload's full size."""
    p = cfg1
    assert p["Yh"].size == 29400
    o, data = oracle_chain(p, THETA1, embed=embed)
    c = engine_chain(p, THETA1, precision="fp64", state_mode="compat", embed=embed)
    assert c.M == o.M == ((80, 96, 80) if embed else (128, 128, 128))
    mu, mom = _seeded_state(o, p, c, (1.3,), seed=1)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < 1e-10 and rel(g_e.imag, g_o.imag) < 1e-10
    _check_leapfrog(o, data, c, mu, mom, 0.05, 25, 1e-9, 1e-9)
    # update from the leapfrog's end point, as the engine now holds it
    mu_e, _ = c.get_state()
    o.mu = mu_e.copy()
    rng = np.random.default_rng(2)
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)], uniforms=[0.41], gammas=[1.0])
    mh = o.update(data, 0.05, 25)
    r = c.update(0.05, 25, 0.41, noise=(zr, zi))
    assert abs(r["H0"] - o.initial_energy) < 1e-9 * abs(o.initial_energy)
    assert abs(r["H1"] - o.proposed_energy) < 1e-9 * abs(o.proposed_energy)
    assert abs(r["mh_rate"] - mh) < 1e-6 and bool(r["accepted"]) == bool(o.accepted)
    mu_e, _ = c.get_state()
    assert rel(mu_e.real, o.mu.real) < 1e-9 and rel(mu_e.imag, o.mu.imag) < 1e-9
    c.close()


# ---------------------------------------------------------------- 1 mm dual resolution

EMBED_1MM = (256, 320, 256)


@pytest.fixture(scope="module")
def mm1():
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    p = problem.dual_problem((128, 160, 128), 1.0, THETA2, seed_gp=20261202, seed_noise=(1, 2))
    W, info = kriging_build(p["hr"], p["q_hr"], p["std"], p["q_std"], THETA2, 6.9, return_info=True)
    p["W"] = W
    p["unique"] = info["n_unique_solves"]
    return p


@pytest.fixture(scope="module")
def mm1_oracle(mm1):
    o, data = oracle_chain(mm1, THETA2, embed=EMBED_1MM)
    return o, data


def test_kriging_1mm_pattern_bit_exact(gpu_available, mm1):
    """The device W at 1 mm against the oracle's integer neighbourhood search
    (kriging_matrix.h:242-299): row pointers and column indices bit-exact
    over all 125,064 rows; float64 weights of sampled interior and boundary
    rows within 1e-9 of the oracle's dense solves; every row sums to 1."""
    from oracle import kriging
    p = mm1
    W = p["W"]
    rows, P, st = kriging.kriging_pattern(p["hr"], p["std"], p["q_hr"], p["q_std"], THETA2, 6.9)
    assert P.shape[0] == 1357 and st["oob"] == 0
    rp = np.zeros(len(rows) + 1, dtype=np.int64)
    rp[1:] = np.cumsum([r[0].size for r in rows])
    assert W.shape == (125064, 1000560)
    assert np.array_equal(np.asarray(W.indptr, dtype=np.int64), rp)
    cols = np.concatenate([r[0] for r in rows])
    assert np.array_equal(np.asarray(W.indices, dtype=np.int64), cols)
    full = np.flatnonzero(np.diff(rp) == 1357)
    assert full.size == 88665
    rs = np.asarray(W.sum(axis=1)).ravel()
    assert np.max(np.abs(rs - 1.0)) < 1e-12
    K = kriging.perturbation_kernel_distances(P, p["q_hr"], THETA2, dtype=np.float32).astype(np.float64)
    rng = np.random.default_rng(3)
    part = np.flatnonzero(np.diff(rp) < 1357)
    sample = np.concatenate([rng.choice(full, 6, replace=False), rng.choice(part, 24, replace=False)])
    for r in sample:
        col, pidx, kp = rows[r]
        w = np.linalg.solve(K[np.ix_(pidx, pidx)], kp.astype(np.float64))
        w /= w.sum()
        got = W.data[rp[r]:rp[r + 1]]
        assert np.max(np.abs(got - w)) <= 1e-9 * np.max(np.abs(w)), r


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dual_res_1mm_gradient_and_leapfrog(gpu_available, mm1, mm1_oracle, precision):
    """Gradient and a 3-step leapfrog with its energies at the 1 mm shape,
    with the device-built W as the shared input."""
    p = mm1
    o, data = mm1_oracle
    c = engine_chain(p, THETA2, precision=precision, state_mode="compat", embed=EMBED_1MM)
    st = c.w_stats()
    assert st["stencil_layout"] and st["stencil_entries"] == 1357 and st["interior_tiles"] == 880
    mu, mom = _seeded_state(o, p, c, (1.1, 3.7), seed=4)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    eg = rel(g_e.real, g_o.real)
    print(f"1 mm {precision} gradient rel err {eg:.2e}", flush=True)
    assert eg < (1e-10 if precision == "fp64" else 2e-4)
    if precision == "fp64":
        _check_leapfrog(o, data, c, mu, mom, 0.3, 3, 1e-9, 1e-9)
    else:
        _check_leapfrog(o, data, c, mu, mom, 0.3, 3, 2e-3, 1e-4)
    c.close()


def test_dual_res_1mm_stencil_layout_matches_csr(gpu_available, mm1):
    """The 1,357-entry interior stencil (gather W mu, tiled W^T v) and the
    packed boundary rows against plain CSR / transposed CSR kernels at 1 mm:
    same terms, different association order (float32: gradient 2e-6,
    energies 1e-5)."""
    from dualres_amd import Chain
    p = mm1
    outs = []
    for force_csr in (True, False):
        c = Chain(p["ijk"], THETA2, p["q_hr"], n_datasets=2, precision="fp32", state_mode="real", seed=4,
                  embed=EMBED_1MM)
        c.w_layout(force_csr)
        c.set_data(p["Yh"], p["Ys"], p["W"])
        c.add_data_to_mu()
        c.set_sigma_sq_inv([1.1, 3.7])
        st = c.w_stats()
        assert st["stencil_layout"] == (not force_csr)
        rng = np.random.default_rng(5)
        z = (rng.standard_normal(c.N), rng.standard_normal(c.N))
        outs.append((c.gradient(), c.update(0.2, 3, 0.5, noise=z), c.log_likelihood()))
        c.close()
    e = rel(outs[1][0], outs[0][0])
    print(f"1 mm stencil vs CSR gradient rel err {e:.2e}", flush=True)
    assert e < 2e-6
    for k in ("H0", "H1"):
        assert abs(outs[1][1][k] - outs[0][1][k]) <= 1e-5 * abs(outs[0][1][k])
    assert abs(outs[1][2] - outs[0][2]) <= 1e-5 * abs(outs[0][2])


def test_dual_res_1mm_fp32_chain_tracks_fp64(gpu_available, mm1):
    """Float32 gradients with float64 energies and accept against the float64
    chain at 1 mm, driven by the same injected normals, accept uniforms and
    Gibbs draws for 12 iterations of L=25: accept decisions agree and the
    state and its running mean stay within 1e-3 (relative)."""
    from dualres_amd.sampler import McmcSummary
    p = mm1
    chains = {pr: engine_chain(p, THETA2, precision=pr, state_mode="compat", embed=EMBED_1MM, seed=3)
              for pr in ("fp64", "fp32")}
    summ = {}
    for pr, c in chains.items():
        c.add_data_to_mu()
        c.set_sigma_sq_inv([1.0, 4.0])
        summ[pr] = McmcSummary(c.n_out, 2, chain=c)
    rng = np.random.default_rng(21)
    N = chains["fp64"].N
    for it in range(12):
        z = (rng.standard_normal(N), rng.standard_normal(N))
        u = rng.uniform()
        r = {pr: c.update(0.25, 25, u, noise=z) for pr, c in chains.items()}
        assert r["fp32"]["accepted"] == r["fp64"]["accepted"], it
        assert abs(r["fp32"]["H1"] - r["fp64"]["H1"]) < 1e-4 * abs(r["fp64"]["H1"]), it
        g = (rng.gamma(0.5 * p["Yh"].size + 1, 2.0 / r["fp64"]["sse_h"]),
             rng.gamma(0.5 * p["Ys"].size + 1, 2.0 / r["fp64"]["sse_s"]))
        for pr, c in chains.items():
            c.set_sigma_sq_inv(list(g))
            summ[pr].update(None, np.ones(2), 0.0)
    m64, m32 = chains["fp64"].mu(), chains["fp32"].mu()
    e1 = rel(m32, m64)
    e2 = rel(summ["fp32"].mode_mu(), summ["fp64"].mode_mu())
    print(f"1 mm fp32 vs fp64: state {e1:.2e}, running mean {e2:.2e}", flush=True)
    assert e1 < 1e-3 and e2 < 1e-3
    for c in chains.values():
        c.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dual_res_1mm_energy_consistent_at_zero_step(gpu_available, mm1, precision):
    """At eps -> 0 the proposal is the start point, so H1 - H0 -> 0: each
    energy piece of the end point must match the start's.  E0 is summed from
    the fresh momentum's spectrum as the end point's kinetic energy is; summed
    from the normals instead, float32 rounding of the transforms left
    k1 - e0 = +2.7 at this size (about 2e7 modes)."""
    p = mm1
    c = engine_chain(p, THETA2, precision=precision, state_mode="compat", embed=EMBED_1MM, seed=8)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.0, 4.0])
    rng = np.random.default_rng(9)
    for _ in range(2):
        z = (rng.standard_normal(c.N), rng.standard_normal(c.N))
        r = c.update(1e-6, 3, 2.0, noise=z)
        print(f"1 mm {precision} eps 1e-6: k1 - e0 {r['k1'] - r['e0']:+.4f}, H1 - H0 {r['H1'] - r['H0']:+.4f}",
              flush=True)
        assert abs(r["k1"] - r["e0"]) < 0.1
        assert abs(r["H1"] - r["H0"]) < 0.1
        # E0 is -1/2 per torus point and stream, to within sampling error
        assert abs(r["e0"] / c.N + 1.0) < 5.0 / np.sqrt(c.N)
    c.close()


# ---------------------------------------------------------------- other transform lengths

@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("embed,dims", [((320, 98, 160), (40, 30, 50)), ((128, 64, 512), (20, 24, 100))])
def test_embeddings_radix7_and_lengths_160_512(gpu_available, precision, embed, dims):
    """(320, 98, 160): x pass of length 160 and z pass of length 160 (both
    compile-time pipelines; 10 modes per thread in the fused z pass), y length
    98 = 2 7^2 (runtime radix 7), and 98 * 161 = 15,778 z columns, so the last
    16-column tile is partly filled.  (128, 64, 512): a 512-long z axis pruned
    to 100 planes.  Gradient and a 4-step leapfrog with energies."""
    from dualres_amd import problem
    p = problem.single_problem(dims, 2.0, THETA1, seed_gp=31, seed_noise=32)
    o, data = oracle_chain(p, THETA1, embed=embed)
    c = engine_chain(p, THETA1, precision=precision, state_mode="compat", embed=embed)
    assert c.M == embed
    mu, mom = _seeded_state(o, p, c, (1.3,), seed=6)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < (1e-10 if precision == "fp64" else 2e-5)
    assert rel(g_e.imag, g_o.imag) < (1e-10 if precision == "fp64" else 2e-5)
    if precision == "fp64":
        _check_leapfrog(o, data, c, mu, mom, 0.05, 4, 1e-9, 1e-9)
    else:
        _check_leapfrog(o, data, c, mu, mom, 0.05, 4, 2e-4, 1e-4)
    c.close()
