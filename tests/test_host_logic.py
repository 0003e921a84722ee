"""Host logic of the run loop (src/main.rs:164-276) on a fake solver: when
progress lines are printed, their format, and the two abort guards (exit 2).
No GPU."""
import math

from cfd_cavity_amd import run


class FakeSolver:
    def __init__(self, nx=8, ny=16, blow_at=None, nan_at=None):
        self._nx, self._ny = nx, ny
        self.t = 0
        self.calls = []
        self.blow_at, self.nan_at = blow_at, nan_at

    def nx(self):
        return self._nx

    def ny(self):
        return self._ny

    def run(self, n):
        self.calls.append(n)
        self.t += n

    def diagnostics(self):
        max_u = 0.2 if self.blow_at is not None and self.t >= self.blow_at else 0.01
        finite = not (self.nan_at is not None and self.t >= self.nan_at)
        return {"max_uy": 0.0123456, "max_u": max_u, "nusselt_hot": 2.5, "nusselt_cold": 2.49996,
                "total_mass": self._nx * self._ny * 1.0000004, "nu_t_max": 3.6e-5,
                "all_finite": finite, "time_step": self.t}


def test_progress_indices_match_reference_condition():
    for steps, every in [(2500, 1000), (1000, 1000), (1, 1000), (7, 3), (0, 5)]:
        ref = [s for s in range(steps) if s % every == 0 or s + 1 == steps]
        assert run.progress_steps(steps, every) == ref


def test_run_issues_steps_between_progress_points():
    s = FakeSolver()
    res = run.run_uniform(s, 2500, 1000, log=lambda _: None)
    assert res.exit_code == 0
    assert s.calls == [1, 1000, 1000, 499] and s.t == 2500
    assert res.lines[0].startswith("step       0 |") and res.lines[-1].startswith("step    2499 |")


def test_progress_line_format():
    line = run.format_progress(1000, FakeSolver().diagnostics(), 128)
    assert line == ("step    1000 | w_max = 0.01235 | max_u = 0.01000 | Nuc = 2.5000 | "
                    "Nuf = 2.5000 | masse = 1.000000 | nu_t_max = 3.600e-5")


def test_abort_on_max_u_guard():
    s = FakeSolver(blow_at=3000)
    res = run.run_uniform(s, 10_000, 1000, log=lambda _: None)
    assert res.exit_code == 2 and s.t == 3001
    assert res.lines[-1].startswith("AVORT : max_u = 0.20000")
    assert math.isclose(run.MAX_U_ABORT, 0.10)


def test_abort_on_non_finite():
    s = FakeSolver(nan_at=2000)
    res = run.run_uniform(s, 10_000, 1000, log=lambda _: None)
    assert res.exit_code == 2 and s.t == 2001
    assert "DIVERGENCE" in res.lines[-1]
