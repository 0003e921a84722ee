"""Diagnostics with the names of the reference's ``post`` module
(bbrangeo/cfd_cavity src/uniform/post.rs), computed on the device in FP64
by ``cavlbm_diagnostics`` (one fused reduction pass over the state)."""
from __future__ import annotations


def nusselt_hot(s):            # post.rs:68-78
    return s.diagnostics()["nusselt_hot"]


def nusselt_hot_wall_link(s):  # post.rs:86-96
    return s.diagnostics()["nusselt_hot_wall_link"]


def nusselt_cold(s):           # post.rs:100-111
    return s.diagnostics()["nusselt_cold"]


def nusselt_volume(s):         # post.rs:116-132
    return s.diagnostics()["nusselt_volume"]


def max_uy_on_row(s, iy):      # post.rs:137-150
    return s.max_uy_on_row(iy)


def max_uy(s):                 # post.rs:153-155
    return s.diagnostics()["max_uy"]


def y_plus_hot(s):             # post.rs:190-199
    return s.diagnostics()["y_plus_hot"]


def check_finite(s):           # post.rs:406-411
    return s.diagnostics()["all_finite"]
