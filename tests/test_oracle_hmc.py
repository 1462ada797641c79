"""Known answers for the oracle's dual-averaging step size
(include/dualres/HMCParameters.h:259-297; Hoffman & Gelman 2014, Alg. 5 as
modified by the reference).  Values below were worked out by hand from the
published recurrences; they are not recomputed with the oracle's code."""
import math

import pytest

from oracle import hmc


def test_initial_search_doubles_then_stops():
    h = hmc.HMCParameters(burnin=100, n_save=10, starting_eps=0.1)
    assert h.integrator_steps() == 1          # L = 1 until a starting eps is found
    h.update(0.9)                             # A = +1; 0.9 > 2^-1 -> eps doubles
    assert h.eps == pytest.approx(0.2)
    h.update(0.7)                             # 0.7 > 0.5 -> 0.4
    assert h.eps == pytest.approx(0.4)
    h.update(0.3)                             # 0.3 <= 0.5 -> found; target log(10 * 0.4)
    assert h.eps == pytest.approx(0.4)
    assert h.eps_start_found and h.integrator_steps() == 10
    assert h.eps_target == pytest.approx(math.log(4.0))


def test_initial_search_halves_when_rate_low():
    h = hmc.HMCParameters(burnin=100, n_save=10, starting_eps=0.8)
    h.update(0.2)                             # A = -1; 0.2^-1 = 5 > 2 -> eps halves
    assert h.eps == pytest.approx(0.4)
    h.update(0.45)                            # 1/0.45 = 2.22 > 2 -> 0.2
    assert h.eps == pytest.approx(0.2)
    h.update(0.6)                             # 1/0.6 = 1.67 <= 2 -> found
    assert h.eps_start_found and h.eps_target == pytest.approx(math.log(2.0))


def test_dual_averaging_steps_known_answers():
    h = hmc.HMCParameters(burnin=100, n_save=10, starting_eps=0.1, mh_target=0.65, gamma=0.05, t0=10, kappa=0.75)
    for r in (0.9, 0.7, 0.3):
        h.update(r)
    # iteration 3, rate 0.5: h_bar = 0.15 / 13, eps = 4 exp(-sqrt(3) * (0.15/13) / 0.05)
    h.update(0.5)
    assert h.h_bar == pytest.approx(0.15 / 13, rel=1e-14)
    assert h.eps == pytest.approx(2.682073873340143, rel=1e-13)
    # eps_bar = eps^(3^-0.75) (eps_bar starts at 1)
    assert h.eps_bar == pytest.approx(2.682073873340143 ** (3 ** -0.75), rel=1e-13)
    assert h.eps_bar == pytest.approx(1.5415811815734066, rel=1e-13)
    # iteration 4, rate 0.8: h_bar = (13/14)(0.15/13) - 0.15/14 = 0 -> eps = exp(target) = 4
    h.update(0.8)
    assert h.h_bar == pytest.approx(0.0, abs=1e-16)
    assert h.eps == pytest.approx(4.0, rel=1e-14)
    assert h.eps_bar == pytest.approx(4 ** (4 ** -0.75) * 1.5415811815734066 ** (1 - 4 ** -0.75), rel=1e-13)


def test_rate_at_target_keeps_eps_at_ten_times_start():
    h = hmc.HMCParameters(burnin=50, n_save=5, starting_eps=0.3, mh_target=0.65)
    h.update(0.2)   # A = -1; 5 > 2 -> 0.15
    h.update(0.55)  # 1/0.55 = 1.82 <= 2 -> found at 0.15; target = log(1.5)
    for _ in range(20):
        h.update(0.65)
        assert h.eps == pytest.approx(1.5, rel=1e-12)


def test_burnin_end_adopts_eps_bar_and_floor():
    h = hmc.HMCParameters(burnin=6, n_save=3, thin=2, starting_eps=0.1)
    for r in (0.9, 0.7, 0.3, 0.5, 0.8, 0.65):
        h.update(r)
    bar = h.eps_bar
    h.update(0.7)                             # iteration == burnin: eps <- eps_bar, rate counted
    assert h.eps == bar and not h.warmup
    assert h.metropolis_hastings_rate() == pytest.approx(0.7)
    assert h.eps_draw(1.1) == pytest.approx(1.1 * bar)
    assert h.max_iterations() == 12
    assert [i for i in range(13) if (setattr(h, "iteration", i) or h.save_iteration())] == [8, 10, 12]
    # the floor applies in dual averaging (the initial search has none): a
    # collapsing rate drives eps down to 1e-5 and holds it there
    f = hmc.HMCParameters(burnin=1000, n_save=1, starting_eps=0.1)
    f.update(0.3)                             # A = -1: 1/0.3 > 2 -> 0.05
    f.update(0.6)                             # found at 0.05
    for _ in range(60):
        f.update(0.0)
        assert f.eps >= 1e-5
    assert f.eps == 1e-5
