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
