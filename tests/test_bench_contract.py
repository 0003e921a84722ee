"""bench.py's host-side contract, on the CPU: it refuses more than one GPU
with a message and exit status 2 before touching a device, and the figures
its roofline block reads from the committed rocprofv3 summaries (calibrated
HBM traffic and FP64 instruction count of the 4-step kernel) parse and give
the stated per-update traffic and issue time."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout == ""


def test_profile_traffic_of_the_four_step_kernel():
    t = bench.measured_traffic(4)
    assert t is not None
    # 256x1024 nodes, 4 lattice updates each: about 79 B per update, well
    # below one-step-at-a-time's 304 B
    per_update = t / (256 * 1024 * 4)
    assert 70.0 < per_update < 90.0


def test_fp64_issue_bound_arithmetic():
    b = bench.fp64_issue_bound(4, 0.0394)
    assert b is not None
    n = b["wave_instructions_per_launch"]
    # 4 cycles per wave64 FP64 instruction over 1024 SIMDs at the peak clock
    assert b["issue_ms"] == pytest.approx(n * 4 / 1024 / 2.4e6, rel=1e-4)
    assert 0.3 < b["issue_frac"] < 0.6
    # all VALU instructions (FP64, integer, FP32) include the FP64 ones
    assert b["all_valu_wave_instructions_per_launch"] > n
    assert b["issue_frac"] < b["all_valu_issue_frac"] < 0.7
    assert bench.fp64_issue_bound(3, 0.03) is None  # no figure for that kernel


def test_form_names_map_to_storages():
    assert [bench.storage_of(f) for f in ("fp64", "fp32", "fp16_offset", "fp64_2step",
                                          "fp64_1step", "fp32_1step", "fp16_offset_1step")] == \
        ["fp64", "fp32", "fp16_offset", "fp64", "fp64", "fp32", "fp16_offset"]
