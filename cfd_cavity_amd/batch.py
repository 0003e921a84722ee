"""Batched form: many independent cavities advanced by one kernel launch per
step (cavlbm_batch_* in include/cavlbm.h).  Each cavity is reachable as a
``CavitySolver`` view for reads and diagnostics."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import CavlbmError, lib
from .solver import CavitySolver, make_params


class CavityBatch:
    """``specs``: list of dicts with CavitySolver::new arguments (nx, ny, ra,
    pr, cs, pr_t, utarget, seed and optionally alpha_mult, nu_override)."""

    def __init__(self, specs, *, storage="fp32", geq_order=2, device=0):
        L = lib()
        arr = (_lib.Params * len(specs))()
        for i, sp in enumerate(specs):
            arr[i] = make_params(sp["nx"], sp["ny"], sp["ra"], sp.get("pr", 0.71), sp["cs"],
                                 sp.get("pr_t", 0.9), sp.get("utarget", 0.02),
                                 sp.get("seed", 42), 1, sp.get("alpha_mult", 1.0), 1.0,
                                 sp.get("nu_override"), storage=storage, geq_order=geq_order,
                                 device=device)
        h = C.c_void_p()
        rc = L.cavlbm_batch_create(arr, len(specs), C.byref(h))
        if rc != 0:
            msg = L.cavlbm_batch_last_error(None) or L.cavlbm_last_error(None)
            raise CavlbmError(rc, msg.decode() if msg else "")
        self._h, self._L = h, L
        self.specs = list(specs)
        self.storage = storage
        self.cavities = []
        for i, sp in enumerate(specs):
            sh = C.c_void_p()
            self._check(L.cavlbm_batch_solver(h, i, C.byref(sh)))
            self.cavities.append(CavitySolver._borrow(sh, sp["nx"], sp["ny"], storage, self))

    def _check(self, rc):
        if rc != 0:
            msg = self._L.cavlbm_batch_last_error(self._h)
            raise CavlbmError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            for c in self.cavities:
                c._h = None
            self._L.cavlbm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.cavities)

    def __getitem__(self, i):
        return self.cavities[i]

    def run(self, nsteps):
        self._check(self._L.cavlbm_batch_run(self._h, int(nsteps)))

    def nodes(self):
        v = C.c_uint64()
        self._check(self._L.cavlbm_batch_nodes(self._h, C.byref(v)))
        return v.value

    def autotune(self, trial_steps=200):
        r = _lib.TuneResult()
        self._check(self._L.cavlbm_batch_autotune(self._h, int(trial_steps), C.byref(r)))
        return {name: getattr(r, name) for name, _ in r._fields_}

    def time_run(self, nsteps):
        ms = C.c_double()
        self._check(self._L.cavlbm_batch_time_run(self._h, int(nsteps), C.byref(ms)))
        return ms.value
