"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must reproduce them bit for bit (CPU), and the HIP path must match
them bit for bit (GPU)."""
import json
import os

import numpy as np
import pytest

from tests.golden.make_golden import CASES, FIELDS, LOG_CASES, LONG_CASES, state_digest
from tests.helpers import bits_equal

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_small():
    return np.load(os.path.join(HERE, "oracle_16x40_210.npz"), allow_pickle=False)


def load_digests():
    with open(os.path.join(HERE, "digests.json")) as fh:
        return json.load(fh)


def test_oracle_reproduces_small_fixture(oracle):
    z = load_small()
    s = oracle.OracleSolver(16, 40, 1.0e5, 0.71, 0.1, 0.9, 0.02, 42, 1)
    s.run(210)
    assert bits_equal(s.f, z["f"]) and bits_equal(s.g, z["g"])
    for n in FIELDS:
        assert bits_equal(getattr(s, n), z[n]), n


def test_oracle_reproduces_digests(oracle):
    dig = load_digests()
    name = "64x256_ra1e6_1000"
    args, steps = CASES[name]
    o = oracle.OracleSolver(*args, 8)
    o.run(steps)
    assert state_digest(o) == dig[name]["sha256"]


LOG_KEYS = ("nusselt_hot", "nusselt_cold", "total_mass", "total_f_mass", "y_plus_hot", "max_u")


def test_oracle_reproduces_log_start(oracle):
    """The logged 64x256 Ra=1e6 run: the oracle reproduces its first three
    log points exactly (the whole run is checked on the GPU)."""
    name = "64x256_ra1e6_20000_log1000"
    rec = load_digests()[name]
    args, steps, every = LOG_CASES[name]
    o = oracle.OracleSolver(*args, 8)
    for ref in rec["log"][:3]:
        o.run(every)
        d = o.diagnostics()
        assert d["time_step"] == ref["step"]
        for k in LOG_KEYS:
            assert d[k] == ref[k], (ref["step"], k)


@pytest.mark.gpu
def test_gpu_matches_small_fixture():
    from cfd_cavity_amd import CavitySolver

    z = load_small()
    s = CavitySolver(16, 40, 1.0e5, 0.71, 0.1, 0.9, 0.02, 42, 1)
    s.run(210)
    assert bits_equal(s.get_pops("f"), z["f"]) and bits_equal(s.get_pops("g"), z["g"])
    for n in FIELDS:
        assert bits_equal(s.get_field(n), z[n]), n


def device_digest(s):
    """sha256 of a device solver's state in the oracle's layout and order."""

    class View:
        pass

    v = View()
    v.f, v.g = s.get_pops("f"), s.get_pops("g")
    for n in FIELDS:
        setattr(v, n, s.get_field(n))
    return state_digest(v)


def assert_diagnostics_close(d, ref, keys=("nusselt_hot", "nusselt_cold", "total_mass",
                                           "y_plus_hot")):
    for k in keys:
        assert abs(d[k] - ref[k]) <= 1e-12 * abs(ref[k]), (k, d[k], ref[k])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES) + sorted(LONG_CASES))
def test_gpu_matches_digests(name):
    """The FP64 state after 1000 steps (64x256 Ra=1e6, 128x512 Ra=1e8), after
    100 000 steps (128x512 Ra=1e8, the drop-in configuration's last
    checkpoint) and after 10 000 steps of 256x1024 Ra=1e10 at utarget 0.005
    (the precision-form configuration) is bitwise the oracle's."""
    from cfd_cavity_amd import CavitySolver

    dig = load_digests()[name]
    args, steps = {**CASES, **LONG_CASES}[name]
    s = CavitySolver(*args, 1)
    s.run(steps)
    assert s.time_step == steps == dig["steps"]
    assert device_digest(s) == dig["sha256"]
    assert_diagnostics_close(s.diagnostics(), dig["diagnostics"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(LOG_CASES))
def test_gpu_matches_oracle_log(name):
    """The reference's own CPU configuration, 64x256 Ra=1e6 for 20 000 steps:
    the diagnostics logged every 1000 steps (Nusselt numbers, both masses,
    y+, max |u|) agree with the oracle's to 1e-12 relative at every point, and
    the final state is bitwise the oracle's."""
    from cfd_cavity_amd import CavitySolver

    rec = load_digests()[name]
    args, steps, every = LOG_CASES[name]
    s = CavitySolver(*args, 1)
    for ref in rec["log"]:
        s.run(every)
        d = s.diagnostics()
        assert d["time_step"] == ref["step"] and d["all_finite"]
        for k in LOG_KEYS:
            assert abs(d[k] - ref[k]) <= 1e-12 * abs(ref[k]), (ref["step"], k, d[k], ref[k])
    assert s.time_step == steps
    assert device_digest(s) == rec["sha256"]
