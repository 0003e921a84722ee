"""Time loop with the reference's progress diagnostics and abort guards
(bbrangeo/cfd_cavity src/main.rs:164-276), driving any solver object with
``run(n)``, ``diagnostics()``, ``nx()``, ``ny()``.

Between two progress points the steps are issued as one ``run(k)`` call, so
the device executes them back to back (graph replays) without host
round-trips; the logged values are those the reference prints after the same
step.
"""
from __future__ import annotations

from dataclasses import dataclass, field

MAX_U_ABORT = 0.10  # src/main.rs:46


@dataclass
class RunResult:
    exit_code: int = 0          # 0 ok, 2 abort (max_u guard or non-finite), main.rs:227/:274
    lines: list = field(default_factory=list)
    last: dict = field(default_factory=dict)


def progress_steps(steps: int, progress_every: int):
    """0-based loop indices at which the reference logs (main.rs:168):
    ``step % progress_every == 0 || step + 1 == steps``; the line for index i is
    printed after i+1 completed steps."""
    out = []
    for i in range(0, steps, progress_every):
        out.append(i)
    if steps > 0 and (steps - 1) not in out:
        out.append(steps - 1)
    return out


def format_progress(step: int, d: dict, n_nodes: int) -> str:
    """The progress line of src/main.rs:187-197."""
    return ("step {:>7} | w_max = {:.5f} | max_u = {:.5f} | Nuc = {:.4f} | Nuf = {:.4f} | "
            "masse = {:.6f} | nu_t_max = {}").format(
        step, d["max_uy"], d["max_u"], d["nusselt_hot"], d["nusselt_cold"],
        d["total_mass"] / n_nodes, _rust_exp(d["nu_t_max"], 3))


def _rust_exp(v: float, prec: int) -> str:
    """Rust's ``{:.3e}`` (mantissa, 'e', exponent without sign padding)."""
    s = f"{v:.{prec}e}"
    mant, exp = s.split("e")
    return f"{mant}e{int(exp)}"


def run_uniform(sim, steps: int, progress_every: int = 1000, log=print) -> RunResult:
    res = RunResult()
    n_nodes = sim.nx() * sim.ny()
    done = 0
    for idx in progress_steps(steps, progress_every):
        sim.run(idx + 1 - done)
        done = idx + 1
        d = sim.diagnostics()
        res.last = d
        line = format_progress(idx, d, n_nodes)
        res.lines.append(line)
        log(line)
        if d["max_u"] > MAX_U_ABORT:
            msg = ("AVORT : max_u = {:.5f} dépasse le garde-fou {:.2f} "
                   "(Ma_local = max_u·√3 ≈ {:.3f} > marge de sécurité)").format(
                d["max_u"], MAX_U_ABORT, d["max_u"] * 3 ** 0.5)
            res.lines.append(msg)
            log(msg)
            res.exit_code = 2
            return res
        if not d["all_finite"]:
            msg = f"DIVERGENCE détectée au pas {idx} (valeurs non finies)."
            res.lines.append(msg)
            log(msg)
            res.exit_code = 2
            return res
    return res
