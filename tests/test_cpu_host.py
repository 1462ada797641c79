"""CPU checks of the host side: library exports, argument validation
without a device, and the sampler's step-size control against the oracle."""
import ctypes
import re

import numpy as np
import pytest

from dualres_amd import _lib as L
from dualres_amd.sampler import HMCParameters
from oracle import hmc as ohmc

HEADER = "include/dualres_hip.h"


def _declared():
    import os
    src = open(os.path.join(os.path.dirname(os.path.dirname(__file__)), HEADER)).read()
    return sorted(set(re.findall(r"\b(dr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = _declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers the declared entry points
    assert set(L.SIGNATURES) | {"dr_last_error"} >= set(names)


def test_host_library_exports_every_declared_symbol():
    import os
    root = os.path.dirname(os.path.dirname(__file__))
    src = open(os.path.join(root, "include", "dualres_host.h")).read()
    names = sorted(set(re.findall(r"\b(drh_[a-z0-9_]+)\s*\(", src)))
    assert len(names) >= 10
    lib = ctypes.CDLL(os.path.join(root, "dualres_amd", "libdualres_host.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_validation_without_device():
    lib = L.load()
    cfg = L.dr_config()
    cfg.n_datasets = 3
    ijk = np.zeros((1, 3), np.int32)
    h = ctypes.c_void_p()
    st = lib.dr_create(ctypes.byref(cfg), L.iptr32(ijk), 1, ctypes.byref(h))
    assert st == L.DR_ELOGIC
    assert b"single or dual" in lib.dr_last_error()
    cfg.n_datasets = 1
    cfg.covfun = L.DR_COV_RQ
    cfg.theta[0], cfg.theta[1], cfg.theta[2] = 1.0, 0.0, 1.0
    assert lib.dr_create(ctypes.byref(cfg), L.iptr32(ijk), 1, ctypes.byref(h)) == L.DR_EDOMAIN
    assert lib.dr_create(None, L.iptr32(ijk), 1, ctypes.byref(h)) == L.DR_ELOGIC
    n = ctypes.c_int64()
    assert lib.dr_mce_offsets(ctypes.byref(n), None) == L.DR_OK
    assert n.value == 25348


@pytest.mark.parametrize("burnin", [0, 5, 40])
def test_hmc_parameters_match_oracle(burnin):
    rng = np.random.default_rng(burnin)
    a = HMCParameters(burnin=burnin, n_save=20, thin=2, integrator_steps=25, starting_eps=0.04)
    b = ohmc.HMCParameters(burnin=burnin, n_save=20, thin=2, integrator_steps=25, starting_eps=0.04)
    for _ in range(burnin + 40):
        mh = float(rng.uniform())
        u = float(rng.uniform(0.9, 1.1))
        assert a.eps(u) == b.eps_draw(u)
        assert a.integrator_steps() == b.integrator_steps()
        assert a.save_iteration() == b.save_iteration()
        a.update(mh)
        b.update(mh)
    assert a.metropolis_hastings_rate() == b.metropolis_hastings_rate()


def test_hmc_parameters_initial_search_and_floor():
    h = ohmc.HMCParameters(burnin=10, starting_eps=0.1)
    assert h.integrator_steps() == 1  # L = 1 until the starting step size is found
    h.update(0.9)  # accept rate > 1/2: double
    assert h.eps == 0.2
    h.update(0.9)
    assert h.eps == 0.4
    h.update(0.1)  # crosses 1/2: search ends
    assert h.eps_start_found and h.integrator_steps() == 10
    for _ in range(6):
        h.update(0.0)
    assert h.eps >= 1e-5
    with pytest.raises(ValueError):
        ohmc.HMCParameters(mh_target=1.0)
