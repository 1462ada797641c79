"""End-to-end runs of the C++ command-line tools on the device: estimate_rbf
(src/estimate_rbf.cpp) and dualgpm/dualgpmf (src/dualgpm.cpp) on small
synthetic NIfTI images."""
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from dualres_amd import kriging, mce, problem
from tests.nifti_io import read_nifti, write_nifti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dualres_amd", "bin")
THETA = (0.806, 0.132, 1.5)


def _write(path, vol_zyx, q):
    write_nifti(str(path), np.transpose(vol_zyx, (2, 1, 0)), pixdim=(q[0, 0], q[1, 1], q[2, 2]),
                qoffset=(q[0, 3], q[1, 3], q[2, 3]))


def _run(args, timeout=120):
    print("cli:", " ".join(os.path.basename(a) if a.startswith(BIN) else a for a in args), flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


@pytest.mark.gpu
def test_estimate_rbf_cli_matches_library(tmp_path, gpu_available):
    p = problem.single_problem((28, 24, 20), 2.0, (1.0, 0.3, 1.5))
    img = tmp_path / "hr.nii.gz"
    _write(img, p["hr"], p["q_hr"])
    csv = tmp_path / "mce.csv"
    r = _run([os.path.join(BIN, "estimate_rbf"), str(img), "--output", str(csv)])
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(csv, delimiter=",", skiprows=1)
    cov, dist, npairs = mce.mce_summary(p["hr"], p["q_hr"])
    assert rows.shape == (len(cov), 3)
    np.testing.assert_allclose(rows[:, 0], dist, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(rows[:, 1], cov, rtol=1e-5, atol=1e-6 * abs(float(cov[0])))
    np.testing.assert_array_equal(rows[:, 2], npairs)
    m = re.search(r"--covariance\s+(\S+)\s+(\S+)\s+(\S+)", r.stdout)
    theta = np.array([float(v) for v in m.groups()])
    assert 0 < theta[0] <= cov[0] * (1 + 1e-5) and theta[1] > 0 and 0 < theta[2] < 2


def _dual_inputs(tmp_path):
    p = problem.dual_problem((24, 20, 16), 1.0, THETA)
    hr, sr = tmp_path / "hr.nii.gz", tmp_path / "sr.nii.gz"
    _write(hr, p["hr"], p["q_hr"])
    _write(sr, p["std"], p["q_std"])
    return p, hr, sr


@pytest.mark.gpu
@pytest.mark.parametrize("tool", ["dualgpmf", "dualgpm"])
def test_dualgpm_cli_dual_resolution(tmp_path, gpu_available, tool):
    p, hr, sr = _dual_inputs(tmp_path)
    base = str(tmp_path / "fit")
    args = [os.path.join(BIN, tool), "--highres", str(hr), "--stdres", str(sr), "--covariance", *map(str, THETA),
            "--burnin", "40", "--nsave", "40", "--thin", "1", "--leapfrog", "10", "--seed", "7", "--samples",
            "--output", base]
    r = _run(args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for suffix in ("_posterior_mean", "_posterior_variance", "_posterior_activation", "_residual"):
        assert os.path.exists(base + suffix + ".nii.gz"), suffix
    mean, q = read_nifti(base + "_posterior_mean.nii.gz")
    var, _ = read_nifti(base + "_posterior_variance.nii.gz")
    res, _ = read_nifti(base + "_residual.nii.gz")
    np.testing.assert_allclose(q, p["q_hr"], atol=1e-6)
    mask = np.transpose(p["mask"], (2, 1, 0))
    y = np.transpose(p["hr"], (2, 1, 0))
    truth = np.transpose(p["truth"], (2, 1, 0))
    assert np.all(np.isfinite(mean[mask])) and np.all(var[mask] > 0)
    assert np.all(mean[~mask] == 0)
    # residual image = data - posterior mean on the data mask (float32 output)
    np.testing.assert_allclose(res[mask], (y - mean)[mask], atol=1e-5 * (1 + np.abs(y[mask]).max()))
    # smoothing the noisy data moves it towards the underlying GP draw
    err_fit = np.sqrt(np.mean((mean[mask] - truth[mask]) ** 2))
    err_data = np.sqrt(np.mean((y[mask] - truth[mask]) ** 2))
    assert err_fit < err_data
    assert os.path.getsize(base + "_samples.dat") > 0
    # the same seed gives the same fit, bit for bit
    base2 = str(tmp_path / "again")
    r2 = _run(args[:-1] + [base2])
    assert r2.returncode == 0
    mean2, _ = read_nifti(base2 + "_posterior_mean.nii.gz")
    np.testing.assert_array_equal(mean, mean2)


@pytest.mark.gpu
def test_dualgpm_cli_single_resolution_with_estimated_covariance(tmp_path, gpu_available):
    p = problem.single_problem((24, 20, 16), 2.0, (1.0, 0.3, 1.5))
    hr = tmp_path / "hr.nii.gz"
    _write(hr, p["hr"], p["q_hr"])
    base = str(tmp_path / "single")
    r = _run([os.path.join(BIN, "dualgpmf"), "--highres", str(hr), "--burnin", "20", "--nsave", "20", "--thin", "1",
              "--leapfrog", "8", "--output", base])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Covariance parameters: (" in r.stdout
    mean, _ = read_nifti(base + "_posterior_mean.nii.gz")
    mask = np.transpose(p["mask"], (2, 1, 0))
    assert np.all(np.isfinite(mean[mask]))


def _datatype(path):
    with open(path, "rb") as f:
        return struct.unpack_from("<h", f.read(72), 70)[0]


def _kriging_inputs(tmp_path, dtype=np.float64):
    p = problem.dual_problem((24, 20, 16), 1.0, THETA)
    src, space = tmp_path / "src.nii.gz", tmp_path / "space.nii"
    write_nifti(str(src), np.transpose(p["hr"], (2, 1, 0)), pixdim=(1.0, 1.0, 1.0),
                qoffset=tuple(p["q_hr"][:3, 3]), dtype=dtype)
    write_nifti(str(space), np.transpose(p["std"], (2, 1, 0)), pixdim=(2.0, 2.0, 2.0),
                qoffset=tuple(p["q_std"][:3, 3]), dtype=dtype)
    fwhm, nu = 6.0, 1.99999
    theta = (1.0, math.log(2.0) / (fwhm / 2) ** nu, nu)  # kernels.inl:45-50
    W = kriging.kriging_build(p["hr"], p["q_hr"], p["std"], p["q_std"], theta, 3 * 2.0)
    return p, src, space, W


@pytest.mark.gpu
def test_write_kriging_matrix_cli(tmp_path, gpu_available):
    p, src, space, W = _kriging_inputs(tmp_path)
    r = subprocess.run([os.path.join(BIN, "write_kriging_matrix"), str(src), "--space", str(space)],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert "W.csv written" in r.stdout
    lines = (tmp_path / "W.csv").read_text().splitlines()
    assert lines[0] == "Data,ColumnIndex,RowCounts,NRow,Ncol"
    rows = [ln.split(",") for ln in lines[1:]]
    assert len(rows) == W.nnz and all(len(f) == 5 for f in rows)
    assert (int(rows[0][3]), int(rows[0][4])) == W.shape
    assert all(f[3] == "" and f[4] == "" for f in rows[1:])
    vals = np.array([float(f[0]) for f in rows])
    np.testing.assert_array_equal([int(f[1]) for f in rows], W.indices)
    nrc = min(W.nnz, W.shape[0] + 1)
    np.testing.assert_array_equal([int(f[2]) for f in rows[:nrc]], W.indptr[:nrc])
    assert all(f[2] == "" for f in rows[nrc:])
    # values printed as float with 6 significant digits
    np.testing.assert_allclose(vals, W.data, rtol=6e-6, atol=1e-30)


@pytest.mark.gpu
def test_interpolate_cli(tmp_path, gpu_available):
    p, src, space, W = _kriging_inputs(tmp_path)
    r = _run([os.path.join(BIN, "interpolate"), str(src), "--space", str(space)])
    assert r.returncode == 0, r.stderr
    out = tmp_path / "src_6mm_fwhm_kriged.nii"
    vol, q = read_nifti(str(out))
    assert _datatype(out) == 64  # the space image's data type
    np.testing.assert_allclose(q, p["q_std"], atol=1e-6)
    sp = np.transpose(p["std"], (2, 1, 0))
    inside = sp != 0
    want = W @ p["Yh"]
    # the space image's nonzero voxels in voxel order (x fastest) receive W Y
    got = np.transpose(vol, (2, 1, 0)).ravel()[np.flatnonzero(p["std"].ravel() != 0)]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.all(vol[~inside] == 0)


@pytest.mark.gpu
def test_interpolate_reproduces_constants(tmp_path, gpu_available):
    """Kriging weights are normalised to sum to one, so a constant image is
    interpolated exactly wherever a row has any neighbour."""
    p, src, space, W = _kriging_inputs(tmp_path, dtype=np.float32)
    const = np.where(p["mask"], 3.25, 0.0)
    write_nifti(str(src), np.transpose(const, (2, 1, 0)), pixdim=(1.0, 1.0, 1.0), qoffset=tuple(p["q_hr"][:3, 3]))
    r = _run([os.path.join(BIN, "interpolate"), str(src), "--space", str(space), "--fwhm", "4.5"])
    assert r.returncode == 0, r.stderr
    vol, _ = read_nifti(str(tmp_path / "src_4mm_fwhm_kriged.nii"))
    assert _datatype(tmp_path / "src_4mm_fwhm_kriged.nii") == 16
    got = np.transpose(vol, (2, 1, 0)).ravel()[np.flatnonzero(p["std"].ravel() != 0)]
    has = np.diff(W.indptr) > 0
    assert has.sum() > 0.9 * len(has)
    np.testing.assert_allclose(got[has], 3.25, rtol=2e-7)
    assert np.all(got[~has] == 0)


@pytest.mark.gpu
def test_dualgpm_cli_update_tau(tmp_path, gpu_available):
    """--update-tau compiles in the reference's tau^-2 Gibbs step
    (MultiResParameters.h:723-725): the tau column of the samples file moves;
    without the flag it stays at sqrt(theta0)."""
    p, hr, sr = _dual_inputs(tmp_path)
    taus = {}
    for flag in (False, True):
        base = str(tmp_path / f"fit_{int(flag)}")
        args = [os.path.join(BIN, "dualgpmf"), "--highres", str(hr), "--stdres", str(sr), "--covariance",
                *map(str, THETA), "--burnin", "20", "--nsave", "8", "--thin", "1", "--leapfrog", "8", "--seed", "3",
                "--samples", "--output", base] + (["--update-tau"] if flag else [])
        r = _run(args)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        rows = [ln.split("\t") for ln in open(base + "_samples.dat").read().splitlines() if ln.strip()]
        assert len(rows) == 8
        taus[flag] = np.array([float(row[-2]) for row in rows])
    np.testing.assert_allclose(taus[False], np.sqrt(THETA[0]), rtol=1e-6)
    assert np.all(taus[True] > 0) and np.all(np.isfinite(taus[True]))
    assert np.ptp(taus[True]) > 0
