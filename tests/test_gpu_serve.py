"""The fit server (dualres_amd/host/server.hpp): whole fits served side by side
on one device must give the same posterior moments, bit for bit, as the same
requests served alone; admission keeps the device bytes under the budget."""
import numpy as np
import pytest

from dualres_amd import problem, serve

THETA = (0.806, 0.132, 1.5)


def _requests():
    s1 = problem.single_problem((16, 14, 12), 2.0, (1.0, 0.3, 1.5), seed_gp=11)
    s2 = problem.single_problem((20, 16, 12), 2.0, (0.9, 0.2, 1.2), seed_gp=12)
    d1 = problem.dual_problem((24, 20, 16), 1.0, THETA, seed_gp=13)
    d2 = problem.dual_problem((20, 20, 16), 1.0, THETA, seed_gp=14)
    kw = dict(burnin=20, nsave=15, thin=1, L=8)
    return [
        serve.make_request(s1["hr"], s1["q_hr"], (1.0, 0.3, 1.5), seed=1, precision="fp32", **kw),
        serve.make_request(d1["hr"], d1["q_hr"], THETA, std=d1["std"], q_std=d1["q_std"], seed=2, precision="fp32",
                           **kw),
        serve.make_request(s2["hr"], s2["q_hr"], (0.9, 0.2, 1.2), seed=3, precision="fp64", **kw),
        serve.make_request(d2["hr"], d2["q_hr"], THETA, std=d2["std"], q_std=d2["q_std"], seed=4, precision="fp64",
                           state_mode="real", **kw),
    ]


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix(gpu_available):
    reqs = _requests()
    budget = 4 * max(serve.estimate_bytes(r) for r in reqs)
    with serve.Server(budget, workers=3) as srv:
        ids = [srv.submit(r) for r in reqs]
        mixed = [srv.wait(i, timeout=300) for i in ids]
    with serve.Server(budget, workers=1) as srv:
        alone = [srv.wait(srv.submit(r), timeout=300) for r in reqs]
    for m, a, r in zip(mixed, alone, reqs):
        assert m["status"] == 0 and a["status"] == 0, (m["error"], a["error"])
        assert m["n_out"] == serve.n_outputs(r)
        assert np.array_equal(m["mean"], a["mean"])
        assert np.array_equal(m["var"], a["var"])
        assert m["mh_rate"] == a["mh_rate"] and m["eps"] == a["eps"]
        assert np.all(np.isfinite(m["mean"])) and np.all(m["var"] >= 0)
        assert m["grad_evals"] == 15 * 9


@pytest.mark.gpu
def test_admission_keeps_budget(gpu_available):
    reqs = [_requests()[1]] * 3
    need = serve.estimate_bytes(reqs[0])
    with serve.Server(int(1.5 * need), workers=3, backfill=False) as srv:
        ids = [srv.submit(r) for r in reqs]
        res = [srv.wait(i, timeout=300) for i in ids]
    assert all(r["status"] == 0 for r in res)
    assert all(np.array_equal(r["mean"], res[0]["mean"]) for r in res)
    # with room for one fit at a time the runs cannot overlap (the three
    # submissions are ~simultaneous, so queue time ~ start time)
    starts = [r["queue_s"] for r in res]
    ends = [r["queue_s"] + r["run_s"] for r in res]
    order = np.argsort(starts)
    for a, b in zip(order[:-1], order[1:]):
        assert starts[b] >= ends[a] - 0.05


@pytest.mark.gpu
def test_request_over_budget_is_refused(gpu_available):
    r = _requests()[1]
    with serve.Server(1024, workers=1) as srv:
        out = srv.wait(srv.submit(r), timeout=30)
    assert out["status"] == 5 and "budget" in out["error"]


@pytest.mark.gpu
def test_device_budget_alone_limits_concurrency(gpu_available):
    """workers <= 0: the server starts a thread for every request its device
    budget admits.  With room for two of three equal fits, two run side by
    side, the third waits for one of them, and all three give the same
    result.  The fits are long enough (a few hundred iterations) for the
    overlap to be seen on the submit clock."""
    reqs = [dict(_requests()[1], burnin=300, nsave=200)] * 3
    need = serve.estimate_bytes(reqs[0])
    with serve.Server(int(2.5 * need), workers=0, backfill=False) as srv:
        ids = [srv.submit(r) for r in reqs]
        res = [srv.wait(i, timeout=300) for i in ids]
    assert all(r["status"] == 0 for r in res)
    assert all(np.array_equal(r["mean"], res[0]["mean"]) for r in res)
    starts = np.array([r["queue_s"] for r in res])
    ends = starts + np.array([r["run_s"] for r in res])
    order = np.argsort(starts)
    first, second, third = order
    assert ends[first] - starts[first] > 0.1
    assert starts[second] < ends[first] - 0.05            # the first two overlap
    assert starts[third] >= min(ends[first], ends[second]) - 0.01


@pytest.mark.gpu
def test_cancel_queued_and_running(gpu_available):
    """A queued request is dropped at once (cancel returns 0); a running fit
    stops at its next HMC iteration (cancel returns 1); both report
    DR_ECANCELED, and the server goes on serving.  Unknown or finished ids
    return 2."""
    import time
    base = _requests()[0]
    long_fit = dict(base, burnin=200000, nsave=15)
    with serve.Server(4 * serve.estimate_bytes(base), workers=1) as srv:
        a = srv.submit(long_fit)
        b = srv.submit(base)
        time.sleep(1.0)  # a is running on the one worker, b waits behind it
        assert srv.cancel(b) == 0
        t0 = time.perf_counter()
        assert srv.cancel(a) == 1
        ra = srv.wait(a, timeout=60)
        stop_s = time.perf_counter() - t0
        rb = srv.wait(b, timeout=10)
        assert ra["status"] == 7 and "cancel" in ra["error"] and ra["mean"] is None
        assert rb["status"] == 7 and rb["run_s"] == 0.0
        assert stop_s < 5.0 and ra["run_s"] >= 0.9
        c = srv.submit(base)
        rc = srv.wait(c, timeout=120)
        assert rc["status"] == 0 and np.all(np.isfinite(rc["mean"]))
        assert srv.cancel(c) == 2 and srv.cancel(12345) == 2
