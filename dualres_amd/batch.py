"""Many chains of uneven grid size advanced together (dr_batch_* of the C ABI).

``Batch(chains, mode)`` borrows created chains; ``update(eps, L, u)`` runs
one HMC iteration of every chain and returns their results.  Modes:
"graph" (every chain's captured iteration a branch of one graph, one launch
per iteration), "streams" (each chain's own graph on its own stream) and
"eager" (plain kernel launches).  A chain's results are bitwise the same in
every mode and alone.
"""
import ctypes as C

import numpy as np

from . import _lib as L

MODES = {"graph": L.DR_BATCH_GRAPH, "streams": L.DR_BATCH_STREAMS, "eager": L.DR_BATCH_EAGER}


class Batch:
    def __init__(self, chains, mode="graph"):
        self._lib = L.load()
        self.chains = list(chains)
        self.mode = mode
        arr = (C.c_void_p * len(self.chains))(*[c._h.value for c in self.chains])
        h = C.c_void_p()
        L.check(self._lib.dr_batch_create(arr, len(self.chains), MODES[mode], C.byref(h)))
        self._h = h

    def update(self, eps, L_steps, u_accept):
        n = len(self.chains)
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float64), (n,)))
        l = np.ascontiguousarray(np.broadcast_to(np.asarray(L_steps, dtype=np.int32), (n,)))
        u = np.ascontiguousarray(np.broadcast_to(np.asarray(u_accept, dtype=np.float64), (n,)))
        res = (L.dr_update_result * n)()
        L.check(self._lib.dr_batch_update(self._h, L.dptr(e), l.ctypes.data_as(C.POINTER(C.c_int)), L.dptr(u), res))
        return [r.as_dict() for r in res]

    def close(self):
        if getattr(self, "_h", None):
            L.check(self._lib.dr_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def fit_batch(chains, hmcs, tau2s, n_s, seeds, mode="graph", timer=None):
    """``sampler.fit`` for many prepared chains in lockstep through one Batch
    (gaussian_process_model.h:134-171 per chain).  Each iteration takes, per
    chain and in fit()'s order, the eps jitter and accept uniform, then one
    batch update of all chains, then each chain's dual averaging, Gibbs draws
    and save.  A chain therefore draws the same numbers and reaches the same
    state as under fit() alone.  The chains' HMCParameters must share burn-in,
    n_save and thin.  ``timer(iteration, timed)`` is called before each
    iteration.  Returns the summaries, per-chain stats and the post-burn-in
    wall time and gradient evaluations of the whole batch."""
    import time

    from .sampler import McmcSummary, gibbs_sigma, init_sigma
    n = len(chains)
    if len({(h.burnin, h.n_save, h.thin) for h in hmcs}) != 1:
        raise ValueError("fit_batch: chains must share burnin, n_save and thin")
    rngs = [np.random.default_rng(s) for s in seeds]
    for c, g, t in zip(chains, rngs, tau2s):
        init_sigma(c, g, t)
    summ = [McmcSummary(c.n_out, c.n_datasets, chain=c) for c in chains]
    saves = [0] * n
    grads = 0
    t_start = None
    h0 = hmcs[0]
    with Batch(chains, mode) as bt:
        while saves[0] < h0.n_save and h0.iteration <= h0.max_iterations():
            timed = h0.iteration >= h0.burnin
            if timer:
                timer(h0.iteration, timed)
            if timed and t_start is None:
                t_start = time.perf_counter()
            eps = np.empty(n)
            Ls = np.empty(n, dtype=np.int32)
            us = np.empty(n)
            for i, (h, g) in enumerate(zip(hmcs, rngs)):
                eps[i] = h.eps(g.uniform(0.9, 1.1))
                Ls[i] = h.integrator_steps()
                us[i] = g.uniform()
            res = bt.update(eps, Ls, us)
            if timed:
                grads += int(np.sum(Ls + 1))
            for i, (c, h, g, r) in enumerate(zip(chains, hmcs, rngs, res)):
                h.update(r["mh_rate"])
                s = gibbs_sigma(c, g, c.n_h, n_s[i], r["sse_h"], r["sse_s"])
                if h.save_iteration():
                    ll = -0.5 * (s[0] * r["sse_h"] + (s[1] * r["sse_s"] if c.n_datasets == 2 else 0.0))
                    summ[i].update(None, np.sqrt(1.0 / np.asarray(s)), ll)
                    saves[i] += 1
    wall = time.perf_counter() - t_start if t_start is not None else 0.0
    stats = [{"mh_rate": h.metropolis_hastings_rate(), "eps": h.eps_value()} for h in hmcs]
    return summ, stats, {"wall_s": wall, "post_burnin_grad_evals": grads}
