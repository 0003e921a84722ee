"""Gradients of the window-averaged hot-wall Nusselt number with respect to
the hot-wall temperature profile and the Smagorinsky constant, through the
hand-written adjoint kernels of libcavlbm (cavlbm_adjoint_nu_window).

    J(theta, Cs) = (1/K) sum_{m=1..K} nusselt_hot(S_m)

over K steps from the solver's current state S_0, which is left unchanged.

``NuWindow`` is a thin ``torch.autograd.Function`` around the C-ABI, so that
the gradient can be checked with torch's tooling (finite differences,
``torch.autograd.gradcheck``-style comparisons); the computation itself runs
in the HIP kernels, the torch tensors live on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _theta(solver, theta):
    if theta is None:
        return None
    a = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    if a.size != solver.ny():
        raise ValueError(f"theta_hot needs {solver.ny()} values, got {a.size}")
    return a


def nu_window(solver, theta, cs, window_steps):
    """J only (forward window; solver state unchanged)."""
    th = _theta(solver, theta)
    j = C.c_double()
    solver._check(solver._L.cavlbm_nu_window(
        solver._h, None if th is None else _dp(th), solver.ny(), float(cs), int(window_steps),
        C.byref(j)))
    return j.value


def nu_window_static(solver, theta, cs, window_steps):
    """J as ``nu_window`` computes it, with the window's buffers in the
    library's static device array instead of device allocations (lattices of
    at most 65536 row-padded nodes and 1024 rows)."""
    th = _theta(solver, theta)
    j = C.c_double()
    solver._check(solver._L.cavlbm_nu_window_static(
        solver._h, None if th is None else _dp(th), solver.ny(), float(cs), int(window_steps),
        C.byref(j)))
    return j.value


def adjoint_nu_window(solver, theta, cs, window_steps, checkpoint_every=100):
    """(J, dJ/dtheta (ny,), dJ/dCs) by the adjoint with checkpoint-and-recompute."""
    th = _theta(solver, theta)
    j = C.c_double()
    gcs = C.c_double()
    gth = np.zeros(solver.ny())
    solver._check(solver._L.cavlbm_adjoint_nu_window(
        solver._h, None if th is None else _dp(th), solver.ny(), float(cs), int(window_steps),
        int(checkpoint_every), C.byref(j), _dp(gth), C.byref(gcs)))
    return j.value, gth, gcs.value


def tangent_nu_window(solver, theta, cs, dtheta, dcs, window_steps):
    """(J, dJ[v]) along v = (dtheta, dcs) by forward-mode FP64 steps."""
    th = _theta(solver, theta)
    dth = _theta(solver, dtheta)
    j = C.c_double()
    dj = C.c_double()
    solver._check(solver._L.cavlbm_tangent_nu_window(
        solver._h, None if th is None else _dp(th), solver.ny(), float(cs),
        None if dth is None else _dp(dth), float(dcs), int(window_steps), C.byref(j),
        C.byref(dj)))
    return j.value, dj.value


def _torch_function():
    import torch

    class NuWindow(torch.autograd.Function):
        """J(theta, cs) with its adjoint gradient; float64 CPU tensors."""

        @staticmethod
        def forward(ctx, theta, cs, solver, window_steps, checkpoint_every):
            j, gth, gcs = adjoint_nu_window(solver, theta.detach().cpu().numpy(),
                                            float(cs.detach().cpu()), window_steps,
                                            checkpoint_every)
            ctx.save_for_backward(torch.from_numpy(gth), torch.tensor(gcs, dtype=torch.float64))
            return torch.tensor(j, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_j):
            gth, gcs = ctx.saved_tensors
            return grad_j * gth, grad_j * gcs, None, None, None

    return NuWindow


def NuWindow():
    """The autograd Function class (built lazily so that torch stays optional)."""
    return _torch_function()
