"""CPU checks of the C++ host layer: NIfTI-1 reader, the RBF minimum-contrast
fit (estimate_kernel_parameters.h:261-387) and the device-free CLIs."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import mce as omce
from oracle import spectrum as ospec
from tests.nifti_io import quaternion_matrix, write_nifti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dualres_amd", "bin")


def _host():
    path = os.path.join(ROOT, "dualres_amd", "libdualres_host.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    f32, i32, f64 = C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    lib.drh_rbf_fit.argtypes = [f32, f32, i32, C.c_int64, f64, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double]
    lib.drh_rbf_fit.restype = C.c_int
    lib.drh_rbf_objective.argtypes = [f32, f32, i32, C.c_int64, f64, C.c_double]
    lib.drh_rbf_objective.restype = C.c_double
    lib.drh_read_nifti.argtypes = [C.c_char_p, C.POINTER(C.c_int), f32, f64]
    lib.drh_read_nifti.restype = C.c_int
    return lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _fit(cov, dist, npairs, theta0, constrained=False, maxdist=1e6, variance=-1.0, bandwidth=-1.0, exponent=-1.0,
         xtol=1e-5):
    lib = _host()
    th = np.array(theta0, dtype=np.float64)
    rc = lib.drh_rbf_fit(_p(cov, C.c_float), _p(dist, C.c_float), _p(npairs, C.c_int32), len(cov),
                         _p(th, C.c_double), int(constrained), maxdist, variance, bandwidth, exponent, xtol)
    return rc, th


def _mce_like(theta, vox=2.0, noise=0.0, seed=0):
    """Empirical-covariance summary shaped like estimate_rbf's: the 25,348
    offsets at isotropic `vox` mm, covariances from the RBF (+ noise)."""
    P = omce.perturbation_matrix_3d()
    dist = omce.offset_distances(P, np.diag([vox, vox, vox, 1.0]))
    dist = dist.astype(np.float32)  # row 0 is the zero offset
    x = dist.astype(np.float64)
    cov = theta[0] * np.exp(-theta[1] * np.power(x, theta[2]))
    rng = np.random.default_rng(seed)
    cov = (cov + noise * rng.standard_normal(cov.shape)).astype(np.float32)
    npairs = rng.integers(50, 5000, size=cov.shape).astype(np.int32)
    return cov, dist, npairs


def test_objective_matches_oracle():
    cov, dist, npairs = _mce_like((1.3, 0.4, 1.4), noise=0.05, seed=3)
    npairs[5::17] = 0  # offsets with no pairs are dropped
    lib = _host()
    for th in [(1.0, 0.6, 1.5), (1.2, 0.3, 1.9), (0.5, 2.0, 0.7)]:
        t = np.array(th)
        got = lib.drh_rbf_objective(_p(cov, C.c_float), _p(dist, C.c_float), _p(npairs, C.c_int32), len(cov),
                                    _p(t, C.c_double), 12.0)
        x, y, w = omce.lsq_data(cov, dist, npairs, maxdist=12.0)
        assert got == pytest.approx(omce.rbf_least_squares(t, x, y, w), rel=1e-12)


def test_fit_recovers_noise_free_parameters():
    truth = (1.3, 0.4, 1.4)
    cov, dist, npairs = _mce_like(truth)
    rc, th = _fit(cov, dist, npairs, (0.8 * cov[0], 0.6, 1.5), xtol=1e-10)
    assert rc == 0
    np.testing.assert_allclose(th, truth, rtol=2e-4)


def test_fit_is_as_good_as_cobyla():
    """With noise the minimiser is not known in closed form: the fit must reach
    an objective no worse than a COBYLA run over the same feasible set."""
    cov, dist, npairs = _mce_like((0.9, 0.25, 1.2), noise=0.03, seed=11)
    theta0 = (0.8 * float(cov[0]), 0.6, 1.5)
    rc, th = _fit(cov, dist, npairs, theta0, xtol=1e-8)
    assert rc == 0
    x, y, w = omce.lsq_data(cov, dist, npairs)
    lb, ub = omce.bounds(float(cov[0]))
    cons = [{"type": "ineq", "fun": (lambda t, k=k: t[k] - lb[k])} for k in range(3)]
    cons += [{"type": "ineq", "fun": (lambda t, k=k: ub[k] - t[k])} for k in (0, 2)]
    ref = minimize(lambda t: omce.rbf_least_squares(t, x, y, w), np.array(theta0), method="COBYLA",
                   constraints=cons, options={"rhobeg": 0.1, "maxiter": 20000, "tol": 1e-12})
    f_ours = omce.rbf_least_squares(th, x, y, w)
    assert all(lb[k] <= th[k] <= ub[k] for k in range(3))
    assert f_ours <= ref.fun * (1 + 1e-6) + 1e-15


def test_fixed_parameters_and_constraint():
    cov, dist, npairs = _mce_like((1.1, 0.3, 1.6), noise=0.01, seed=5)
    rc, th = _fit(cov, dist, npairs, (1.0, 0.6, 1.5), exponent=1.6)
    assert rc == 0 and abs(th[2] - 1.6) <= 1e-5
    rc, th = _fit(cov, dist, npairs, (1.0, 0.6, 1.5), bandwidth=0.3, exponent=1.6)
    assert rc == 0 and abs(th[1] - 0.3) <= 1e-5 and abs(th[2] - 1.6) <= 1e-5
    # the constrained fit keeps psi <= nu
    rc, th = _fit(cov, dist, npairs, (1.0, 0.6, 1.5), constrained=True)
    assert rc == 0 and th[1] <= th[2] + 1e-12
    # a fixed bandwidth above the exponent violates the constraint
    rc, _ = _fit(cov, dist, npairs, (1.0, 0.6, 1.5), constrained=True, bandwidth=1.8, exponent=1.0)
    assert rc == 2
    # all three fixed: returned as given
    rc, th = _fit(cov, dist, npairs, (1.0, 0.6, 1.5), variance=0.5, bandwidth=0.2, exponent=1.0)
    assert rc == 0 and tuple(th) == (0.5, 0.2, 1.0)


@pytest.mark.parametrize("case", ["f32_gz", "f64_plain", "i16_big_endian", "rotated_qfac"])
def test_nifti_reader(tmp_path, case):
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((7, 5, 4))
    kw = dict(pixdim=(1.5, 2.0, 2.5), qoffset=(-10.0, 4.0, 7.5))
    name, dtype = "a.nii", np.float64
    if case == "f32_gz":
        name, dtype = "a.nii.gz", np.float32
    elif case == "i16_big_endian":
        vol = np.round(vol * 1000)
        dtype = np.int16
        kw["big_endian"] = True
    elif case == "rotated_qfac":
        s = math.sqrt(0.5)
        kw.update(quatern=(0.0, 0.0, s), qfac=-1.0)  # 90 degrees about z, left-handed
    path = str(tmp_path / name)
    write_nifti(path, vol, dtype=dtype, **kw)
    lib = _host()
    dims = (C.c_int * 3)()
    q = np.zeros(16, np.float32)
    data = np.zeros(vol.size)
    assert lib.drh_read_nifti(path.encode(), dims, _p(q, C.c_float), _p(data, C.c_double)) == 0
    assert tuple(dims) == vol.shape
    np.testing.assert_array_equal(data, np.asarray(vol, dtype).astype(np.float64).ravel(order="F"))
    Q = q.reshape(4, 4)
    expect = np.eye(4)
    if case == "rotated_qfac":
        # closed form: 90-degree rotation about z maps x->y, y->-x; qfac flips z
        expect[:3, :3] = np.array([[0, -2.0, 0], [1.5, 0, 0], [0, 0, -2.5]])
    else:
        expect[:3, :3] = np.diag([1.5, 2.0, 2.5])
    expect[:3, 3] = (-10.0, 4.0, 7.5)
    np.testing.assert_allclose(Q, expect, atol=1e-6)
    q2 = np.eye(4)
    q2[:3, :3] = quaternion_matrix(*kw.get("quatern", (0, 0, 0)), kw["pixdim"], kw.get("qfac", 1.0))
    q2[:3, 3] = kw["qoffset"]
    np.testing.assert_allclose(Q, q2, atol=1e-6)


def test_nifti_reader_rejects_garbage(tmp_path):
    path = tmp_path / "bad.nii"
    path.write_bytes(b"\x00" * 400)
    lib = _host()
    dims = (C.c_int * 3)()
    q = np.zeros(16, np.float32)
    assert lib.drh_read_nifti(str(path).encode(), dims, _p(q, C.c_float), None) == 1


def test_rbf_neighborhood_cli():
    exe = os.path.join(BIN, "rbf_neighborhood")
    theta = (0.9, 0.35, 1.6)
    out = subprocess.run([exe, "0.2", "--covariance", *map(str, theta)], capture_output=True, text=True, check=True)
    assert float(out.stdout.split()[0]) == pytest.approx(ospec.rbf_inverse(0.2, theta[1], theta[2], theta[0]), rel=1e-5)
    bad = subprocess.run([exe, "1.5", "--covariance", *map(str, theta)], capture_output=True, text=True)
    assert bad.returncode == 1 and "between (0, 1)" in bad.stderr


def test_cli_usage_errors():
    for name in ("dualgpm", "dualgpmf", "estimate_rbf"):
        exe = os.path.join(BIN, name)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 1 and "Usage" in (r.stdout + r.stderr)
        h = subprocess.run([exe, "--help"], capture_output=True, text=True)
        assert h.returncode == 0
    r = subprocess.run([os.path.join(BIN, "dualgpm"), "--highres", "/nonexistent.nii"], capture_output=True, text=True)
    assert r.returncode == 1


def test_kriging_cli_usage_errors(tmp_path):
    """KrigingCommandParser.inl:66-176: usage, help, missing inputs, bad values."""
    img = tmp_path / "img.nii"
    write_nifti(str(img), np.ones((4, 4, 4), np.float32))
    for name in ("interpolate", "write_kriging_matrix"):
        exe = os.path.join(BIN, name)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 1 and "Usage" in r.stderr
        assert subprocess.run([exe, "--help"], capture_output=True, text=True).returncode == 0
        r = subprocess.run([exe, str(img)], capture_output=True, text=True)
        assert r.returncode == 1 and "must supply a --space" in r.stderr
        r = subprocess.run([exe, "--space", str(img)], capture_output=True, text=True)
        assert r.returncode == 1 and "must supply a NIfTI file input" in r.stderr
        r = subprocess.run([exe, str(img), "--space", str(img), "--exponent", "2.5"], capture_output=True, text=True)
        assert r.returncode == 1 and "bad parameter value (exponent" in r.stderr
        r = subprocess.run([exe, str(img), "--space", str(tmp_path / "missing.nii")], capture_output=True, text=True)
        assert r.returncode == 1 and "not a NIfTI file" in r.stderr
        r = subprocess.run([exe, str(img), "--bogus"], capture_output=True, text=True)
        assert r.returncode == 1 and "Unrecognized option" in r.stderr


def _hmc_trace(rates, burnin=100, n_save=10, thin=1, L=10, starting_eps=0.1, mh_target=0.65, gamma=0.05, t0=10.0,
               kappa=0.75):
    """Runs the C++ HMCParameters<double> of the CLIs (host/sampler.hpp) over a
    rate sequence; one row per update: eps, eps_bar, h_bar, eps_target, L,
    eps_start_found, warmup, mh rate."""
    lib = _host()
    lib.drh_hmc_trace.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 5 + \
        [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    lib.drh_hmc_trace.restype = C.c_int
    r = np.ascontiguousarray(rates, dtype=np.float64)
    out = np.zeros((len(r), 8))
    rc = lib.drh_hmc_trace(burnin, n_save, thin, L, starting_eps, mh_target, gamma, t0, kappa, _p(r, C.c_double),
                           len(r), _p(out, C.c_double))
    assert rc == 0
    return out


def test_cpp_hmc_parameters_initial_search():
    """HMCParameters.h:259-279 in the C++ class: the starting eps doubles while
    the rate stays above 1/2 (A = +1), halves while it stays below (A = -1);
    L = 1 until found; then target = log(10 eps).  Hand-worked values, the
    same known answers the oracle's HMCParameters is held to."""
    t = _hmc_trace([0.9, 0.7, 0.3], starting_eps=0.1)
    assert t[:, 0] == pytest.approx([0.2, 0.4, 0.4])
    assert list(t[:, 4]) == [1, 1, 10] and list(t[:, 5]) == [0, 0, 1]
    assert t[2, 3] == pytest.approx(math.log(4.0))
    t = _hmc_trace([0.2, 0.45, 0.6], starting_eps=0.8)
    assert t[:, 0] == pytest.approx([0.4, 0.2, 0.2])
    assert t[2, 3] == pytest.approx(math.log(2.0)) and t[2, 5] == 1


def test_cpp_hmc_parameters_dual_averaging_known_answers():
    """Dual averaging (HMCParameters.h:280-290; Hoffman & Gelman 2014 Alg. 5):
    after the search ends at eps = 0.4 (iteration 2), iteration 3 at rate 0.5
    gives h_bar = 0.15/13 and eps = 4 exp(-sqrt(3) h_bar / 0.05) =
    2.682073873340143, eps_bar = eps^(3^-0.75) = 1.5415811815734066; rate 0.8
    at iteration 4 returns h_bar to 0 and eps to exp(target) = 4."""
    t = _hmc_trace([0.9, 0.7, 0.3, 0.5, 0.8])
    assert t[3, 2] == pytest.approx(0.15 / 13, rel=1e-14)
    assert t[3, 0] == pytest.approx(2.682073873340143, rel=1e-13)
    assert t[3, 1] == pytest.approx(1.5415811815734066, rel=1e-13)
    assert t[4, 2] == pytest.approx(0.0, abs=1e-16)
    assert t[4, 0] == pytest.approx(4.0, rel=1e-14)
    assert t[4, 1] == pytest.approx(4 ** (4 ** -0.75) * 1.5415811815734066 ** (1 - 4 ** -0.75), rel=1e-13)
    # a rate at the target keeps eps at ten times the start
    t = _hmc_trace([0.2, 0.55] + [0.65] * 20, burnin=50, n_save=5, starting_eps=0.3)
    assert t[2:, 0] == pytest.approx(np.full(20, 1.5), rel=1e-12)


def test_cpp_hmc_parameters_burnin_end_and_floor():
    """At iteration == burnin the sampler adopts eps_bar and starts counting MH
    rates (HMCParameters.h:259-266, 292-296); during dual averaging eps never
    falls below 1e-5 (HMCParameters.h:287)."""
    rates = [0.9, 0.7, 0.3, 0.5, 0.8, 0.65, 0.7, 0.5]
    t = _hmc_trace(rates, burnin=6, n_save=3, thin=2)
    assert t[6, 0] == t[5, 1] and t[6, 6] == 0 and t[5, 6] == 1
    assert t[6, 7] == pytest.approx(0.7) and t[7, 7] == pytest.approx(0.6)
    t = _hmc_trace([0.3, 0.6] + [0.0] * 60, burnin=1000, n_save=1)
    assert np.all(t[2:, 0] >= 1e-5) and t[-1, 0] == 1e-5
