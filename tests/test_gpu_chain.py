"""GPU parity of the HMC chain against the CPU oracle (float64 and float32).

The oracle (oracle/mrp.py) restates MultiResParameters.h with complex state
and c2c FFTs; the engine runs a Fourier-domain real state with pruned
r2c/c2r passes.  Both are driven from the same injected state and noise.

Tolerances (relative 2-norm against the float64 oracle):
  float64 engine: eigenvalues 1e-12 of max|lambda|, gradient 1e-10,
                  leapfrog trajectory 1e-9, energies 1e-9 (relative);
  float32 engine: gradient 2e-4, trajectory 2e-3, energies 1e-4.
"""
import numpy as np
import pytest

from helpers import engine_chain, oracle_chain, random_state, rel, small_dual, small_single
from oracle import mrp

pytestmark = pytest.mark.gpu

THETA1 = (0.806, 0.132, 1.0)
THETA2 = (0.806, 0.132, 1.5)


def _setup(p, theta, precision, mode, embed=None, sig=(1.3, 0.7), seed=0):
    o, data = oracle_chain(p, theta, embed=embed)
    c = engine_chain(p, theta, precision=precision, state_mode=mode, embed=embed)
    rng = np.random.default_rng(seed)
    mu, mom = random_state(o, rng, imag=(mode == "compat"))
    mu.reshape(-1)[o.I] += p["Yh"]
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(s) for s in sig[: o.n_datasets]]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv(list(sig[: o.n_datasets]))
    return o, data, c, mu, mom


@pytest.mark.parametrize("embed", [None, (20, 24, 18)])
def test_eigenvalues_fp64(gpu_available, embed):
    p = small_single()
    o, _ = oracle_chain(p, THETA1, embed=embed)
    c = engine_chain(p, THETA1, embed=embed)
    assert c.M == o.M
    lam_e = c.lambda_()
    lam_o = o.lam.real
    assert np.max(np.abs(lam_e - lam_o)) <= 1e-12 * np.max(np.abs(lam_o))
    assert c.n_nonpositive == o.Z.size
    assert abs(c.max_eig - lam_o.ravel()[0] / o.N) <= 1e-12 * abs(lam_o.ravel()[0] / o.N)


@pytest.mark.parametrize("embed", [None, (20, 24, 18)])
def test_state_roundtrip(gpu_available, embed):
    p = small_single()
    c = engine_chain(p, THETA1, embed=embed)
    rng = np.random.default_rng(1)
    mu = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    pm = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    c.set_state(mu, pm)
    mu2, pm2 = c.get_state()
    assert rel(mu2, mu) < 1e-13
    assert rel(pm2, pm) < 1e-13


@pytest.mark.parametrize("mode", ["compat", "real"])
@pytest.mark.parametrize("embed", [None, (20, 24, 18)])
def test_gradient_single_fp64(gpu_available, mode, embed):
    p = small_single()
    o, data, c, mu, _ = _setup(p, THETA1, "fp64", mode, embed)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < 1e-10
    if mode == "compat":
        assert rel(g_e.imag, g_o.imag) < 1e-10


@pytest.mark.parametrize("mode", ["compat", "real"])
def test_gradient_dual_fp64(gpu_available, mode):
    p = small_dual()
    o, data, c, mu, _ = _setup(p, THETA2, "fp64", mode)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < 1e-10
    # energies of the injected state
    assert abs(c.log_prior() - o.log_prior()) < 1e-9 * abs(o.log_prior())
    assert abs(c.log_likelihood() - o.log_likelihood(data)) < 1e-9 * abs(o.log_likelihood(data))


@pytest.mark.parametrize("dual", [False, True])
@pytest.mark.parametrize("mode", ["compat", "real"])
def test_leapfrog_fp64(gpu_available, dual, mode):
    p = small_dual() if dual else small_single()
    theta = THETA2 if dual else THETA1
    o, data, c, mu, mom = _setup(p, theta, "fp64", mode)
    o._compute_mass_matrix_eigen_values()
    eps, L = 0.07, 6
    lp0 = o._log_posterior(data, mu)
    k0 = o._potential_energy(mom)
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), eps, L)
    lp1 = o._log_posterior(data, mu1)
    k1 = o._potential_energy(p1)
    r = c.leapfrog(eps, L)
    mu_e, p_e = c.get_state()
    assert rel(mu_e.real, mu1.real) < 1e-9
    assert rel(p_e.real, p1.real) < 1e-9
    if mode == "compat":
        assert rel(mu_e.imag, mu1.imag) < 1e-9
    assert abs((r["log_prior1"] + r["log_lik1"]) - lp1) < 1e-9 * abs(lp1)
    assert abs((r["log_prior0"] + r["log_lik0"]) - lp0) < 1e-9 * abs(lp0)
    assert abs(r["k1"] - k1) < 1e-9 * abs(k1)
    assert abs(r["e0"] - k0) < 1e-9 * abs(k0)


@pytest.mark.parametrize("dual", [False, True])
def test_update_injected_noise_fp64(gpu_available, dual):
    """Full iteration (momentum draw, leapfrog, energies, accept) vs the oracle."""
    p = small_dual() if dual else small_single()
    theta = THETA2 if dual else THETA1
    o, data, c, mu, _ = _setup(p, theta, "fp64", "compat")
    rng = np.random.default_rng(42)
    zr, zi = rng.standard_normal(o.N), rng.standard_normal(o.N)
    eps, L = 0.05, 5
    u = 0.37
    o.rng = mrp.ScriptedRng(normals=[(zr, zi)], uniforms=[u], gammas=[1.0, 1.0])
    mh_o = o.update(data, eps, L)
    r = c.update(eps, L, u, noise=(zr.reshape(o.shape), zi.reshape(o.shape)))
    assert abs(r["H0"] - o.initial_energy) < 1e-9 * abs(o.initial_energy)
    assert abs(r["H1"] - o.proposed_energy) < 1e-9 * abs(o.proposed_energy)
    assert abs(r["mh_rate"] - mh_o) < 1e-6
    assert bool(r["accepted"]) == bool(o.accepted)
    mu_e, _ = c.get_state()
    assert rel(mu_e.real, o.mu.real) < 1e-9
    assert rel(mu_e.imag, o.mu.imag) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_nan_energy_is_rejected(gpu_available, precision):
    """A proposal whose energy is NaN has R = NaN, which the reference turns
    into R = 0 (MultiResParameters.h:714): it is rejected even with a zero
    accept uniform, mh_rate is 0, the state is untouched, and the next
    iteration is the one a chain that never saw the NaN would run.  The NaN
    enters through one injected momentum normal."""
    p = small_dual()
    o, data, c, mu, _ = _setup(p, THETA2, precision, "compat")
    rng = np.random.default_rng(8)
    zr, zi = rng.standard_normal(o.N).reshape(o.shape), rng.standard_normal(o.N).reshape(o.shape)
    bad = zr.copy()
    bad.reshape(-1)[5] = np.nan
    o.rng = mrp.ScriptedRng(normals=[(bad.reshape(-1), zi.reshape(-1))], uniforms=[0.0], gammas=[1.0, 1.0])
    with np.errstate(invalid="ignore"):
        mh_o = o.update(data, 0.05, 4)
    assert mh_o == 0.0 and not o.accepted
    mu0, _ = c.get_state()
    r = c.update(0.05, 4, 0.0, noise=(bad, zi))
    assert np.isnan(r["H0"]) or np.isnan(r["H1"])
    assert r["mh_rate"] == 0.0 and not r["accepted"]
    mu1, _ = c.get_state()
    assert np.array_equal(mu1, mu0)
    # the next iteration equals that of a twin chain set to the same state
    _, _, twin, _, _ = _setup(p, THETA2, precision, "compat")
    twin.set_sigma_sq_inv(c.sigma_sq_inv())
    a = c.update(0.05, 4, 0.5, noise=(zr, zi))
    b = twin.update(0.05, 4, 0.5, noise=(zr, zi))
    assert a["H0"] == b["H0"] and a["H1"] == b["H1"] and a["accepted"] == b["accepted"]
    assert np.array_equal(c.get_state()[0], twin.get_state()[0])
    c.close()
    twin.close()


@pytest.mark.parametrize("dual", [False, True])
def test_gradient_leapfrog_fp32(gpu_available, dual):
    p = small_dual() if dual else small_single()
    theta = THETA2 if dual else THETA1
    o, data, c, mu, mom = _setup(p, theta, "fp32", "compat")
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < 2e-4
    o._compute_mass_matrix_eigen_values()
    eps, L = 0.07, 6
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), eps, L)
    lp1 = o._log_posterior(data, mu1)
    r = c.leapfrog(eps, L)
    mu_e, p_e = c.get_state()
    assert rel(mu_e.real, mu1.real) < 2e-3
    assert rel(p_e.real, p1.real) < 2e-3
    assert abs((r["log_prior1"] + r["log_lik1"]) - lp1) < 1e-4 * abs(lp1)


def test_output_gather_and_init(gpu_available):
    """add_data_to_mu + mu_() gather: mu_ == mu at data indices; init draw scale."""
    p = small_single()
    c = engine_chain(p, THETA1)
    mu0 = c.mu()
    c.add_data_to_mu()
    mu1 = c.mu()
    assert rel(mu1 - mu0, p["Yh"]) < 1e-12
    c.set_output_indices(p["ijk"])
    assert np.allclose(c.mu_(), mu1, rtol=0, atol=1e-12)


def test_too_many_negative_eigenvalues(gpu_available):
    """A constant kernel (psi = 0) has N-1 zero eigenvalues on a power-of-two
    torus, so the constructor raises EDOMAIN (MultiResParameters.h:283-284)."""
    from dualres_amd import DrDomainError
    p = small_single()
    with pytest.raises(DrDomainError):
        engine_chain(p, (1.0, 0.0, 1.0))


@pytest.mark.parametrize("geom", ["aligned", "oblique"])
def test_kriging_build_matches_oracle(gpu_available, geom):
    """Pattern (row pointers, columns) bit-exact; float64 weights within 1e-6."""
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    from oracle import kriging
    theta = (0.806, 0.132, 1.5)
    p = problem.dual_problem((14, 12, 10), 2.0, theta, seed_gp=9)
    q_std = p["q_std"].copy()
    if geom == "oblique":
        q_std[0, 3] += 0.37  # stdres origin off the half-voxel grid
        q_std[1, 3] -= 0.21
    radius = 2.3 * 2.0
    rp, cols, vals, nrow, ncol, st = kriging.get_sparse_kriging_matrix_data(
        p["hr"], p["std"], p["q_hr"], q_std, theta, radius)
    W, info = kriging_build(p["hr"], p["q_hr"], p["std"], q_std, theta, radius, return_info=True)
    assert W.shape == (nrow, ncol)
    assert np.array_equal(W.indptr, rp)
    assert np.array_equal(W.indices, cols)
    assert np.max(np.abs(W.data - vals)) <= 1e-6 * max(1.0, np.max(np.abs(vals)))
    assert info["n_unique_solves"] <= nrow


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dims", [(16, 14, 12), (24, 20, 16)])
def test_w_stencil_layout_matches_csr(gpu_available, precision, dims):
    """Stencil + boundary W kernels against the plain CSR kernels: same terms,
    different association order only (tolerances: fp64 1e-13; fp32 2e-6 on the
    gradient, 1e-5 on energies after three leapfrog steps)."""
    from dualres_amd import Chain
    p = small_dual(dims=dims)
    outs = []
    for force_csr in (True, False):
        c = Chain(p["ijk"], THETA2, p["q_hr"], n_datasets=2, precision=precision, state_mode="real", seed=4)
        c.w_layout(force_csr)
        c.set_data(p["Yh"], p["Ys"], p["W"])
        c.set_sigma_sq_inv([1.1, 0.6])
        st = c.w_stats()
        assert st["stencil_layout"] == (not force_csr)
        if not force_csr:
            assert st["interior_rows"] > 0 and st["boundary_rows"] > 0
            assert st["interior_rows"] + st["boundary_rows"] == p["W"].shape[0]
            assert st["stencil_entries"] * st["interior_rows"] + st["boundary_entries"] == p["W"].nnz
            assert st["interior_tiles"] > 0 and st["tile_cells"] * (4 if precision == "fp32" else 8) <= 48 * 1024
        rng = np.random.default_rng(5)
        z = (rng.standard_normal(c.N), rng.standard_normal(c.N))
        outs.append((c.gradient(), c.update(0.05, 3, 0.5, noise=z), c.log_likelihood()))
        c.close()
    tol = 1e-13 if precision == "fp64" else 2e-6
    assert rel(outs[1][0], outs[0][0]) < tol
    htol = 1e-13 if precision == "fp64" else 1e-5
    for k in ("H0", "H1"):
        assert abs(outs[1][1][k] - outs[0][1][k]) <= htol * abs(outs[0][1][k])
    assert abs(outs[1][2] - outs[0][2]) <= htol * abs(outs[0][2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_summary_matches_host_moments(gpu_available, precision):
    """mcmc_summary::update on the device against the same moments built on
    the host from mu_() (gaussian_process_model.h:205-224); both float64."""
    from dualres_amd.sampler import McmcSummary
    p = small_dual(dims=(16, 14, 12))
    c = engine_chain(p, THETA2, precision=precision, state_mode="real")
    c.add_data_to_mu()
    out = p["ijk"][::3]
    c.set_output_indices(out)
    c.set_sigma_sq_inv([1.1, 0.6])
    host = McmcSummary(c.n_out, 2)
    dev = McmcSummary(c.n_out, 2, chain=c)
    rng = np.random.default_rng(9)
    for k in range(4):
        z = (rng.standard_normal(c.N), rng.standard_normal(c.N))
        c.update(0.05, 3, 0.3, noise=z)
        host.update(c.mu_(), np.ones(2), -1.0)
        dev.update(None, np.ones(2), -1.0)
    for a, b in ((dev.m1, host.m1), (dev.m2, host.m2), (dev.m1_abs, host.m1_abs)):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-300)
    # var = (m2 - m1^2/n)/n cancels: the derived activation keeps ~1e-10 of the moments' 1e-14
    np.testing.assert_allclose(dev.activation(), host.activation(), rtol=1e-9)
    assert c.summary_read()[3] == 4
    c.set_output_indices(out)          # resets the running sums
    assert c.summary_read()[3] == 0 and not c.summary_read()[0].any()
    c.close()


def test_fp32_chain_tracks_fp64(gpu_available):
    """FP32 gradients and FFTs with FP64 energies and accept: driven by the
    same injected normals, uniforms and Gibbs draws, the FP32 chain stays within
    1e-3 (relative) of the FP64 chain over 40 dual-resolution HMC iterations,
    posterior running mean included."""
    from dualres_amd.sampler import McmcSummary
    p = small_dual(dims=(24, 20, 16))
    chains = {pr: engine_chain(p, THETA2, precision=pr, state_mode="compat", seed=3) for pr in ("fp64", "fp32")}
    summ = {}
    for c in chains.values():
        c.add_data_to_mu()
        c.set_output_indices(p["ijk"])
        c.set_sigma_sq_inv([1.3, 0.7])
    for pr, c in chains.items():
        summ[pr] = McmcSummary(c.n_out, 2, chain=c)
    rng = np.random.default_rng(21)
    for it in range(40):
        z = (rng.standard_normal(chains["fp64"].N), rng.standard_normal(chains["fp64"].N))
        u = rng.uniform()
        r = {pr: c.update(0.08, 6, u, noise=z) for pr, c in chains.items()}
        assert r["fp32"]["accepted"] == r["fp64"]["accepted"], it
        g = (rng.gamma(20.0, 0.06), rng.gamma(10.0, 0.07))
        for pr, c in chains.items():
            c.set_sigma_sq_inv(list(g))
            summ[pr].update(None, np.ones(2), 0.0)
    m64, m32 = chains["fp64"].mu(), chains["fp32"].mu()
    assert rel(m32, m64) < 1e-3
    assert rel(summ["fp32"].mode_mu(), summ["fp64"].mode_mu()) < 1e-3
    for c in chains.values():
        c.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tau_update_matches_oracle(gpu_available, precision):
    """dr_set_tau_sq_inv rescales lambda as _update_tau_sq_inv does
    (MultiResParameters.h:870-871); eigenvalues, gradient and prior then match
    the oracle after the same assignment.  dr_mu_data_sumsq is the rate input
    |Re mu at I|^2 (:866-867).  Repeated assignments rescale from the base
    spectrum, so returning to 1/theta0 restores lambda bit for bit."""
    p = small_dual()
    o, data, c, mu, _ = _setup(p, THETA2, precision, "compat")
    lam0 = c.lambda_()
    assert c.tau_sq_inv() == pytest.approx(1 / THETA2[0], rel=1e-15)
    s = c.mu_data_sumsq()
    assert s == pytest.approx(float(np.sum(c.mu() ** 2)), rel=1e-12 if precision == "fp64" else 1e-6)
    assert s == pytest.approx(float(np.sum(o._gather_re(o.mu) ** 2)), rel=1e-12 if precision == "fp64" else 1e-6)
    for t in (0.4, 2.9):
        c.set_tau_sq_inv(t)
        o.set_tau_sq_inv(t)
        tol = 1e-12 if precision == "fp64" else 1e-6
        assert rel(c.lambda_(), o.lam.real) < tol
        g_e, g_o = c.gradient(), o.compute_gradient(data, o.mu)
        assert rel(g_e.real, g_o.real) < (1e-10 if precision == "fp64" else 2e-4)
        assert c.log_prior() == pytest.approx(o.log_prior(), rel=1e-9 if precision == "fp64" else 1e-4)
    c.set_tau_sq_inv(1 / THETA2[0])
    assert np.array_equal(c.lambda_(), lam0)
    with pytest.raises(ValueError):
        c.set_tau_sq_inv(-1.0)
    c.close()


def test_fit_with_tau_update(gpu_available):
    """fit(update_tau=True): the tau^-2 draw follows the sigma draws each
    iteration; the chain stays finite and tau moves away from theta0."""
    from dualres_amd.sampler import HMCParameters, fit
    p = small_dual(dims=(16, 14, 12))
    c = engine_chain(p, THETA2, precision="fp32", state_mode="compat", seed=5)
    c.add_data_to_mu()
    c.set_output_indices(p["ijk"])
    hmc = HMCParameters(burnin=10, n_save=10, thin=1, integrator_steps=6, starting_eps=0.1, mh_target=0.65)
    summ, st = fit(c, hmc, THETA2[0], n_s=p["Ys"].size, seed=4, update_tau=True)
    t = c.tau_sq_inv()
    assert np.isfinite(t) and t > 0 and abs(t - 1 / THETA2[0]) > 1e-6
    assert np.all(np.isfinite(summ.mode_mu()))
    c.close()


def _ratio_problem(dims, ratio, theta, radius, seed=3):
    """Highres image on `dims` at 1 mm and a stdres image at `ratio` mm whose
    voxel centres fall on highres voxels (bases on the lattice ratio * Z^3)."""
    from dualres_amd import problem
    from oracle import kriging as okr
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    m = problem.ellipsoid_mask(dims, 0.45)
    hr = np.where(m, rng.standard_normal(m.shape) + 3.0, 0.0)
    q_hr = problem.qform_axis(1.0, dims)
    sd = tuple(max(2, d // ratio) for d in dims)
    sm = problem.ellipsoid_mask(sd, 0.45)
    std = np.where(sm, rng.standard_normal(sm.shape) + 3.0, 0.0)
    off = (ratio - 1) / 2.0
    q_std = problem.qform_axis(float(ratio), sd, origin=[q_hr[a, 3] + off for a in range(3)])
    rp, cols, vals, nrow, ncol, _ = okr.get_sparse_kriging_matrix_data(hr, std, q_hr, q_std, theta, radius)
    p = {"hr": hr, "q_hr": q_hr, "std": std, "q_std": q_std, "ijk": problem.nonzero_ijk(hr),
         "Yh": hr.ravel()[np.flatnonzero(problem.present(hr).ravel())],
         "Ys": std.ravel()[np.flatnonzero(problem.present(std).ravel())]}
    p["W"] = sp.csr_matrix((vals, cols, rp), shape=(nrow, ncol))
    return p


@pytest.mark.parametrize("ratio,dims,precision", [(1, (14, 12, 10), "fp64"), (3, (27, 24, 21), "fp64"),
                                                  (3, (27, 24, 21), "fp32")])
def test_w_stencil_layout_other_lattices(gpu_available, ratio, dims, precision):
    """The stencil layout's tiled interior W^T v and gather W mu on lattices of
    spacing 1 (same-resolution kriging, one residue class) and 3 (27 classes)
    against the plain CSR kernels."""
    from dualres_amd import Chain
    p = _ratio_problem(dims, ratio, THETA2, radius=2.2 * ratio)
    outs = []
    for force_csr in (True, False):
        c = Chain(p["ijk"], THETA2, p["q_hr"], n_datasets=2, precision=precision, state_mode="real", seed=4)
        c.w_layout(force_csr)
        c.set_data(p["Yh"], p["Ys"], p["W"])
        c.set_sigma_sq_inv([1.1, 0.6])
        st = c.w_stats()
        if not force_csr:
            assert st["stencil_layout"] and st["interior_rows"] > 0 and st["interior_tiles"] > 0
            assert st["interior_rows"] + st["boundary_rows"] == p["W"].shape[0]
            assert st["stencil_entries"] * st["interior_rows"] + st["boundary_entries"] == p["W"].nnz
            if ratio == 3:
                assert st["far_rows"] > 0  # neighbourhoods that wrap across a row end
        rng = np.random.default_rng(6)
        z = (rng.standard_normal(c.N), rng.standard_normal(c.N))
        outs.append((c.gradient(), c.update(0.05, 3, 0.5, noise=z), c.log_likelihood()))
        c.close()
    tol = 1e-13 if precision == "fp64" else 2e-6
    assert rel(outs[1][0], outs[0][0]) < tol
    htol = 1e-13 if precision == "fp64" else 1e-5
    for k in ("H0", "H1"):
        assert abs(outs[1][1][k] - outs[0][1][k]) <= htol * abs(outs[0][1][k])
    assert abs(outs[1][2] - outs[0][2]) <= htol * abs(outs[0][2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gradient_long_z_embedding(gpu_available, precision):
    """A 256-long z axis, as in the 1 mm workload, with the z passes pruned to
    40 planes: gradient, a leapfrog trajectory and every energy of it
    (reduced over the fused z pass's and the half kicks' tiles) against the
    oracle."""
    p = small_single(dims=(6, 6, 40))
    embed = (16, 16, 256)
    o, data, c, mu, mom = _setup(p, THETA1, precision, "compat", embed)
    g_o = o.compute_gradient(data, o.mu)
    g_e = c.gradient()
    assert rel(g_e.real, g_o.real) < (1e-10 if precision == "fp64" else 2e-5)
    o._compute_mass_matrix_eigen_values()
    eps, L = 0.05, 4
    lp0 = o._log_posterior(data, mu)
    k0 = o._potential_energy(mom)
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), eps, L)
    lp1 = o._log_posterior(data, mu1)
    k1 = o._potential_energy(p1)
    r = c.leapfrog(eps, L)
    mu_e, p_e = c.get_state()
    tol, etol = (1e-9, 1e-9) if precision == "fp64" else (2e-4, 1e-4)
    assert rel(mu_e.real, mu1.real) < tol
    assert rel(p_e.real, p1.real) < tol
    assert rel(mu_e.imag, mu1.imag) < tol
    assert abs((r["log_prior1"] + r["log_lik1"]) - lp1) < etol * abs(lp1)
    assert abs((r["log_prior0"] + r["log_lik0"]) - lp0) < etol * abs(lp0)
    assert abs(r["k1"] - k1) < etol * abs(k1)
    assert abs(r["e0"] - k0) < etol * abs(k0)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rational_quadratic_spectrum_and_gradient(gpu_available, precision):
    """--rational-quadratic (kernels.inl:58-74, MultiResParameters.h:234-243):
    c = tau^2 (1 + psi/nu b^2)^-nu on the torus base, lambda = F(c).  Device
    eigenvalues (always float64) against the oracle to 1e-12 of max|lambda|,
    the nonpositive count exactly, then gradient and a leapfrog trajectory."""
    from dualres_amd import Chain
    theta = (0.9, 0.4, 1.3)
    p = small_single(dims=(12, 10, 8), theta=theta)
    embed = (24, 20, 18)
    o = mrp.MultiResParameters(1, theta, p["ijk"], p["q_hr"], rng=mrp.NumpyRng(3), covfun="rq", M=embed)
    data = mrp.MultiResData(p["Yh"])
    c = Chain(p["ijk"], theta, p["q_hr"], n_datasets=1, precision=precision, state_mode="compat", covfun="rq",
              embed=embed, seed=3)
    c.set_data(p["Yh"])
    lam_o = o.lam.real
    assert np.max(np.abs(c.lambda_() - lam_o)) <= 1e-12 * np.max(np.abs(lam_o))
    assert c.n_nonpositive == o.Z.size
    rng = np.random.default_rng(8)
    mu, mom = random_state(o, rng)
    mu.reshape(-1)[o.I] += p["Yh"]
    o.mu = mu.copy()
    o.sigma_sq_inv = [np.float64(1.3)]
    c.set_state(mu, mom)
    c.set_sigma_sq_inv([1.3])
    g_o = o.compute_gradient(data, o.mu)
    assert rel(c.gradient().real, g_o.real) < (1e-10 if precision == "fp64" else 2e-5)
    o._compute_mass_matrix_eigen_values()
    mu1, p1 = o.leapfrog(data, mu.copy(), mom.copy(), 0.06, 5)
    c.leapfrog(0.06, 5)
    mu_e, p_e = c.get_state()
    tol = 1e-9 if precision == "fp64" else 2e-4
    assert rel(mu_e.real, mu1.real) < tol and rel(p_e.real, p1.real) < tol
    c.close()


def test_kriging_pattern_flat_offset_wrap(gpu_available):
    """A highres image with every voxel present, so neighbourhoods of rows
    near the x, y and z faces run past the grid.  The reference checks only
    the flat offset (kriging_matrix.h:263-266): a neighbour past the x face
    lands at the other end of the previous or next row and is kept.  Device
    pattern bit-exact against the oracle, with such wrapped entries present."""
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    from oracle import kriging
    theta = (0.806, 0.132, 1.5)
    dims, sd = (14, 12, 10), (7, 6, 5)
    rng = np.random.default_rng(1)
    hr = rng.standard_normal(dims[::-1]) + 3.0
    std = rng.standard_normal(sd[::-1]) + 3.0
    q_hr = problem.qform_axis(2.0, dims)
    q_std = problem.qform_axis(4.0, sd, origin=[q_hr[a, 3] + 1.0 for a in range(3)])
    rp, cols, vals, nrow, ncol, st = kriging.get_sparse_kriging_matrix_data(hr, std, q_hr, q_std, theta, 4.6)
    rows, P, _ = kriging.kriging_pattern(hr, std, q_hr, q_std, theta, 4.6)
    base = 2 * problem.nonzero_ijk(std)       # stdres voxel centres sit on highres voxel 2 ijk
    wrapped = sum(int(np.any((base[i] + P[sel] < 0) | (base[i] + P[sel] >= np.array(dims)), axis=1).sum())
                  for i, (_, sel, _) in enumerate(rows))
    assert wrapped > 1000
    W = kriging_build(hr, q_hr, std, q_std, theta, 4.6)
    assert np.array_equal(W.indptr, rp) and np.array_equal(W.indices, cols)
    assert np.max(np.abs(W.data - vals)) <= 1e-6 * max(1.0, np.max(np.abs(vals)))


@pytest.mark.parametrize("dims,vox", [((14, 12, 10), 2.0), ((18, 16, 14), 1.0)])
def test_kriging_indefinite_systems_tau2_above_one(gpu_available, dims, vox):
    """tau^2 = 1.5 with the reference's unit diagonal (kriging_matrix.h:118)
    makes K indefinite, so the device Cholesky meets a nonpositive pivot and
    those systems are factorised again by the device LU with partial
    pivoting: pattern bit-exact and weights within 1e-6 of the oracle's
    pivoted float64 solves (kriging_matrix.h:309-313 solves the same systems
    by column-pivoted QR).  At 2 mm the systems fit one 64-tile (panel
    factorisation only); at 1 mm they have 437 unknowns, 7 tiles, so the
    blocked triangular solve and the MFMA trailing update run too."""
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    from oracle import kriging
    theta = (1.5, 0.132, 1.5)
    p = problem.dual_problem(dims, vox, theta, seed_gp=19)
    radius = 4.6
    P = kriging.neighborhood_perturbation(p["q_hr"], radius)
    K = kriging.perturbation_kernel_distances(P, p["q_hr"], theta).astype(np.float64)
    assert np.linalg.eigvalsh(K).min() < 0          # indefinite
    rp, cols, vals, nrow, ncol, st = kriging.get_sparse_kriging_matrix_data(
        p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    W, info = kriging_build(p["hr"], p["q_hr"], p["std"], p["q_std"], theta, radius, return_info=True)
    assert info["n_lu"] > 0 and info["n_lu"] + info["n_cholesky"] == info["n_unique_solves"]
    assert np.array_equal(W.indptr, rp) and np.array_equal(W.indices, cols)
    assert np.max(np.abs(W.data - vals)) <= 1e-6 * max(1.0, np.max(np.abs(vals)))
    assert np.allclose(np.asarray(W.sum(axis=1)).ravel()[np.diff(rp) > 0], 1.0, atol=1e-12)


def test_kriging_singular_systems_drop_their_rows(gpu_available):
    """psi = 0 with tau^2 = 1 makes every K (unit diagonal) a matrix of ones:
    systems of two or more neighbours are singular, the Cholesky and then
    the pivoted LU flag them, and their rows come out empty, as rows whose
    weights sum to NaN do (kriging_matrix.h:320-327); rows with a single
    neighbour keep it with weight 1, at the compacted offsets.  At radius
    2.1 mm on this grid: 1 empty, 11 single-neighbour and 70 singular rows."""
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    from oracle import kriging
    p = problem.dual_problem((14, 12, 10), 2.0, (0.806, 0.132, 1.5), seed_gp=9)
    theta, radius = (1.0, 0.0, 1.5), 2.1
    rows, P, st = kriging.kriging_pattern(p["hr"], p["std"], p["q_hr"], p["q_std"], theta, radius)
    cnt = np.array([r[0].size for r in rows])
    assert (cnt == 0).sum() == 1 and (cnt == 1).sum() == 11 and (cnt >= 2).sum() == 70
    W, info = kriging_build(p["hr"], p["q_hr"], p["std"], p["q_std"], theta, radius, return_info=True)
    assert W.shape[0] == len(rows) and info["n_lu"] > 0
    got = np.diff(W.indptr)
    assert np.array_equal(got, np.where(cnt == 1, 1, 0))
    single = np.flatnonzero(cnt == 1)
    assert np.array_equal(W.indices, np.concatenate([rows[r][0] for r in single]))
    assert np.array_equal(W.data, np.ones(single.size))


def test_native_fit_loop_on_a_prepared_chain(gpu_available):
    """The C++ host's fit loop (dualgpmf's fit_model iterations) driven on a
    chain prepared through the C ABI: the same seed reproduces the run bit
    for bit, post_iters bounds the timed iterations, the step size tunes
    during burn-in and the acceptance rate is a probability."""
    from dualres_amd.sampler import fit_native
    p = small_dual(dims=(16, 14, 12))
    outs = []
    for captured in (False, False, True):
        c = engine_chain(p, THETA2, precision="fp32", state_mode="compat", seed=6)
        c.add_data_to_mu()
        c.set_output_indices(p["ijk"])
        r = fit_native(c, p["Ys"].size, THETA2[0], burnin=20, nsave=10, thin=3, L=6, seed=12, post_iters=25,
                       captured=captured)
        outs.append((r, c.mu(), c.summary_read()))
        c.close()
    (r1, m1, s1), (r2, m2, s2), (r3, m3, s3) = outs
    # iterations replayed from captured graphs give the same run bit for bit
    assert r3["mh_rate"] == r1["mh_rate"] and r3["eps"] == r1["eps"] and r3["grad_evals"] == r1["grad_evals"]
    assert np.array_equal(m3, m1) and np.array_equal(s3[0], s1[0]) and np.array_equal(s3[1], s1[1])
    # every timed iteration runs L = 6 steps, except a first one of 1 step
    # when burn-in ends before the step-size search does (as in the reference)
    assert r1["iterations"] == 25 and 25 * 7 - 5 <= r1["grad_evals"] <= 25 * 7
    assert 0.0 < r1["mh_rate"] <= 1.0 and r1["eps"] != 1.0 / 6
    assert r1["mh_rate"] == r2["mh_rate"] and r1["eps"] == r2["eps"]
    assert np.array_equal(m1, m2) and np.array_equal(s1[0], s2[0]) and s1[3] == 8
