"""The host layer's CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/host_sanitize.cpp is compiled with g++ -fsanitize=address,undefined
together with dualres_amd/host/nifti.cpp and run on NIfTI inputs written
here: a rotated anisotropic volume with a mask and NaNs, a file cut inside
its header, and a gzip file whose data stops short of what its header
declares.  Any sanitizer report or failed check fails the test."""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.nifti_io import write_nifti

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _build(out):
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", os.path.join(HERE, "host_sanitize.cpp"),
           os.path.join(ROOT, "dualres_amd", "host", "nifti.cpp"), "-lz", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "asan" in (r.stderr + r.stdout).lower():
        pytest.skip("the sanitizer runtime is not installed")
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_layer_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    _build(exe)
    rng = np.random.default_rng(4)
    vol = rng.standard_normal((11, 9, 7)).astype(np.float32)
    vol[:2] = 0.0
    vol[5, 4, 3] = np.nan
    good = tmp_path / "in.nii.gz"
    write_nifti(str(good), vol, pixdim=(1.8, 1.8, 2.3), qoffset=(-91.3, 112.7, -54.1), quatern=(0.061, -0.034, 0.118))
    plain = tmp_path / "plain.nii"
    write_nifti(str(plain), vol)
    trunc = tmp_path / "trunc.nii"
    trunc.write_bytes(plain.read_bytes()[:200])
    short = tmp_path / "short.nii.gz"
    with gzip.open(short, "wb") as f:
        f.write(plain.read_bytes()[:352 + 4 * 100])
    # verify_asan_link_order=0: a preloaded library may come before the ASan
    # runtime in the process's library list
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(good), str(tmp_path), str(trunc), str(short)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-4000:])
