"""Batched form: many independent cavities advanced by one kernel launch per
step (cavlbm_batch_* in include/cavlbm.h).  Each cavity is reachable as a
``CavitySolver`` view for reads and diagnostics."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import CavlbmError, lib
from .solver import CavitySolver, make_params

# Grid sizes of the batched workload, all of aspect ratio Az = 4 as the
# reference's cavities (README.md:69 lists 64x256, 96x384, 128x512; its Ra=1e10
# runs use 256x1024, README.md:167).
BATCH_SIZES = [(64, 256), (96, 384), (128, 512), (192, 768), (256, 1024)]


def seeded_batch_specs(n=64, seed=42):
    """``n`` cavities: sizes drawn (seeded) from BATCH_SIZES, Ra log-uniform in
    [1e6, 1e9], Cs uniform in [0.08, 0.12], Pr 0.71, Pr_t 0.9, utarget 0.02,
    cavity i seeded with seed + i (numpy PCG64, stable across versions)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        nx, ny = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        ra = float(10 ** rng.uniform(6.0, 9.0))
        cs = float(rng.uniform(0.08, 0.12))
        specs.append(dict(nx=nx, ny=ny, ra=ra, cs=cs, pr=0.71, pr_t=0.9, utarget=0.02,
                          seed=seed + i))
    return specs


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
