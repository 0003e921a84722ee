"""CAVSNAP1 snapshots and the bilinear regrid of restarts, with the names of
the reference's ``snapshot`` module (bbrangeo/cfd_cavity
src/uniform/snapshot.rs:14-131), through libcavlbm's host-side I/O."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import CavlbmError, lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc != 0:
        raise CavlbmError(rc, lib().cavlbm_snapshot_last_error().decode())


def _arrs(*fields):
    return [np.ascontiguousarray(f, dtype=np.float64).ravel() for f in fields]


def write_snapshot(path, nx, ny, rho, ux, uy, t):
    a = _arrs(rho, ux, uy, t)
    _check(lib().cavlbm_snapshot_write(os.fsencode(path), nx, ny, *[_dp(x) for x in a]))


def write_snapshot_atomic(path, nx, ny, rho, ux, uy, t):
    a = _arrs(rho, ux, uy, t)
    _check(lib().cavlbm_snapshot_write_atomic(os.fsencode(path), nx, ny, *[_dp(x) for x in a]))


def read_snapshot(path):
    """(nx, ny, rho, ux, uy, t)"""
    L = lib()
    nx, ny = C.c_uint64(), C.c_uint64()
    _check(L.cavlbm_snapshot_read(os.fsencode(path), C.byref(nx), C.byref(ny), None, None,
                                  None, None, 0))
    n = nx.value * ny.value
    out = [np.empty(n) for _ in range(4)]
    _check(L.cavlbm_snapshot_read(os.fsencode(path), C.byref(nx), C.byref(ny),
                                  *[_dp(x) for x in out], n))
    return (nx.value, ny.value, *out)


def interpolate_field(src, nx0, ny0, nx1, ny1):
    s = _arrs(src)[0]
    dst = np.empty(nx1 * ny1)
    _check(lib().cavlbm_interpolate_field(_dp(s), nx0, ny0, _dp(dst), nx1, ny1))
    return dst


def restart(solver, path):
    """The --restart sequence of src/main.rs:104-128: read, regrid to the
    solver's grid, init_from_fields (f = feq, g = geq, time_step = 0)."""
    snx, sny, rho, ux, uy, t = read_snapshot(path)
    nx, ny = solver.nx(), solver.ny()
    fields = [interpolate_field(a, snx, sny, nx, ny) for a in (rho, ux, uy, t)]
    solver.init_from_fields(*fields)
    return snx, sny


def save(solver, path):
    """write_snapshot_atomic of the solver's rho, ux, uy, T (main.rs:284-300)."""
    write_snapshot_atomic(path, solver.nx(), solver.ny(), solver.get_field("rho"),
                          solver.get_field("ux"), solver.get_field("uy"),
                          solver.get_field("temperature"))


# ---------------------------------------------------------------------------
# OCTSNAP1: multi-level snapshots of the octree solver (src/octree/snapshot.rs)

def write_octree_snapshot(path, levels, step, grids, atomic=True):
    """``grids``: list of (nx, ny, rho, ux, uy, T) in grid order
    (snapshot.rs:57-79, atomic write snapshot.rs:140-161)."""
    dims = (C.c_uint64 * (2 * len(grids)))(*[v for g in grids for v in g[:2]])
    fields = [np.concatenate([np.asarray(g[2 + q], dtype=np.float64).ravel() for g in grids])
              for q in range(4)]
    _check(lib().cavlbm_octsnap_write(os.fsencode(path), levels, step, len(grids), dims,
                                      *[_dp(x) for x in fields], 1 if atomic else 0))


def read_octree_snapshot(path):
    """(levels, step, grids) with grids = [(nx, ny, rho, ux, uy, T), ...]
    (snapshot.rs:83-137, :164-170)."""
    L = lib()
    lv, ng = C.c_uint32(), C.c_uint32()
    st = C.c_uint64()
    dims = (C.c_uint64 * 10)()
    p = os.fsencode(path)
    _check(L.cavlbm_octsnap_read(p, C.byref(lv), C.byref(st), C.byref(ng), dims, None, None,
                                 None, None, 0))
    d = [(dims[2 * g], dims[2 * g + 1]) for g in range(ng.value)]
    total = sum(nx * ny for nx, ny in d)
    out = [np.empty(total) for _ in range(4)]
    _check(L.cavlbm_octsnap_read(p, C.byref(lv), C.byref(st), C.byref(ng), dims,
                                 *[_dp(x) for x in out], total))
    grids, off = [], 0
    for nx, ny in d:
        n = nx * ny
        grids.append((nx, ny) + tuple(a[off:off + n] for a in out))
        off += n
    return lv.value, st.value, grids
