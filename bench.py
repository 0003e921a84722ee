"""Benchmark of the MI355X cavity time step: MLUPS and fraction of the HBM
roofline at 256x1024, plus the FP32 / FP16-offset storage forms against FP64.

Workload (one GPU): D2Q9 f+g double-distribution cavity 256x1024, Ra=1e10,
Pr=0.71, Pr_t=0.9, Cs=0.10, utarget=0.005, seed 42, the reference's init
(linear T + xorshift64* noise of +-1e-3).  A step is one fused
stream+BC+moments+collide pass over the whole lattice; inputs stay resident
in HBM.  The run loop advances several steps per kernel launch (temporal
blocking, bitwise equal to single steps: 4 in FP64, 3 in FP32 and FP16-offset); the
"fp64_2step" and "fp64_1step" forms time two steps per launch and the
single-step kernel beside it.  The timed region (exactly --steps steps after
--warmup untimed ones) runs eleven times; ms_per_step and value come from HIP
events on the solver stream (median of the eleven: a 20-step region lasts
about 0.2 ms and single regions vary by a few percent), the host wall clock
around the synchronisations is reported beside them.  The GPU's clock settles
only under sustained load (the same launch takes the same cycles but up to
15 % longer right after an idle gap), so each form is measured twice on fresh
solvers: right after set-up (reported as *_after_idle) and right after
--gpu-load-ms of sustained load on a throwaway solver of the same workload
(the reported figures, the clocks of a long run).  Prints ONE JSON line.

The CPU baseline is the oracle (a C restatement of the reference's row-band
rayon path) timed on this host's cores on a bounded sample of the same
workload.  The other paths (batched cavities, the dynamic SGS step, the
octree, the adjoint, the run loop's time averages and snapshots) are timed
beside the headline under "batched" and "other_paths".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic bytes per lattice update: read + write 9 f and 9 g populations
# and the per-node Guo force fy_node (FP16-offset: plus the FP32 per-node
# temperature plane its g offsets refer to).  A launch of the N-step kernel
# processes N lattice updates per node, so its algorithmic bytes are
# N * nodes * these, while it moves the state across HBM once (the measured
# traffic, from the committed counter summaries).
BYTES_PER_LUP = {"fp64": 2 * (18 * 8 + 8), "fp32": 2 * (18 * 4 + 4),
                 "fp16_offset": 2 * (18 * 2 + 4 + 4)}
DTYPE = {"fp64": "f64", "fp32": "f32", "fp16_offset": "f16-offset storage, f32 arithmetic"}
REFERENCE_CPU_MLUPS_256x1024 = 49.8  # README/PLAN_GPU_BACKEND: 2M steps in ~2.9 h, jobs 8, M1 Max
METRIC = "MLUPS & %HBM peak at 256x1024; mixed vs FP64 at rel dNu<=1e-4"

WORKLOAD = dict(nx=256, ny=1024, ra=1.0e10, pr=0.71, cs=0.10, pr_t=0.9, utarget=0.005, seed=42)


def storage_of(form: str) -> str:
    """The storage of a form name ("fp32", "fp64_2step" -> "fp64", ...)."""
    if form.endswith("step") and "_" in form:
        return form.rsplit("_", 1)[0]
    return form


def make_solver(storage: str, tblock=None):
    from cfd_cavity_amd import CavitySolver

    w = WORKLOAD
    s = CavitySolver(w["nx"], w["ny"], w["ra"], w["pr"], w["cs"], w["pr_t"], w["utarget"],
                     w["seed"], 1, storage=storage)
    if tblock is not None:
        s.time_blocking = tblock
    return s


def timed_regions(s, steps: int, reps: int):
    """`reps` repetitions of exactly `steps` steps, each bracketed by stream
    and device synchronisations and timed by HIP events recorded on the
    solver's stream around its launches and by the host clock; returns the
    per-repetition event and wall times (ms)."""
    import torch

    ev, walls = [], []
    for _ in range(reps):
        s.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev.append(s.time_run(steps))
        s.synchronize()
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)
    return ev, walls


def load_gpu(storage: str, tblock, ms: float):
    """Keeps the GPU busy for about `ms` milliseconds with the same workload
    on a throwaway solver.  The GPU's clocks settle only under sustained load:
    the same 4-step launch, at the same cycle count, takes 39-42 us after an
    idle gap and 36.2 us after about 100 ms of load, whatever the flow state
    (profiles/k_stepn_clock_state_256x1024.md).  Returns the solver, which the
    caller destroys after its measurement (a destruction in between would
    leave the GPU idle again), and the steps it ran."""
    if ms <= 0:
        return None, 0
    w = make_solver(storage, tblock)
    t0 = time.perf_counter()
    n = 0
    while (time.perf_counter() - t0) * 1e3 < ms:
        w.run(400)
        w.synchronize()
        n += 400
    return w, n


def gpu_form(form: str, steps: int, warmup: int, kernel_launches: int, reps: int = 11,
             gpu_load_ms: float = 300.0):
    """One storage form with the library's default steps per launch; a name
    ending in "_<n>step" runs the same storage with at most n steps per
    launch (n = 1: the single-step kernel), for comparison.

    After `warmup` untimed steps the timed region -- exactly `steps` steps --
    runs `reps` times in a row (timed_regions); the figures are the medians
    over the repetitions.  This is done twice on fresh solvers of the
    workload: right after the solver is set up (the GPU coming out of idle,
    reported as *_after_idle), and then, for the reported figures, right
    after `gpu_load_ms` of sustained load on a throwaway solver
    (load_gpu), which brings the GPU to the clocks a long run sees; the
    measured solver's own warmup and timed steps are unchanged."""
    import statistics

    storage, tblock = storage_of(form), None
    if storage != form:
        tblock = int(form.rsplit("_", 1)[1][:-len("step")])
    nodes = WORKLOAD["nx"] * WORKLOAD["ny"]
    cold = make_solver(storage, tblock)
    cold.run(warmup)
    ev_cold, _ = timed_regions(cold, steps, reps)
    cold.close()
    s = make_solver(storage, tblock)
    spl = s.steps_per_launch
    load, load_steps = load_gpu(storage, tblock, gpu_load_ms)
    s.run(warmup)
    ev, walls = timed_regions(s, steps, reps)
    if load is not None:
        load.close()
    ev_ms = statistics.median(ev)
    wall = statistics.median(walls) * 1e-3
    d = s.diagnostics()
    kern_ms = s.time_step_kernel(kernel_launches) if kernel_launches > 0 else float("nan")
    mlups = nodes * steps / (ev_ms * 1e-3) / 1e6
    b = BYTES_PER_LUP[storage]
    # average step-kernel launch duration over the timed region, from HIP
    # events recorded on the solver's stream (steps / spl launches), and the
    # algorithmic bytes of the spl lattice updates per node of one launch
    launch_ms = ev_ms / (steps / spl)
    achieved = spl * nodes * b / (launch_ms * 1e-3) / 1e9
    res = {
        "steps_per_launch": spl,
        "mlups": mlups,
        "ms_per_step": ev_ms / steps,
        "ms_per_step_wall": wall * 1e3 / steps,
        "mlups_wall": nodes * steps / wall / 1e6,
        "ms_per_step_events_reps": [round(v / steps, 6) for v in ev],
        "ms_per_step_after_idle": statistics.median(ev_cold) / steps,
        "mlups_after_idle": nodes * steps / (statistics.median(ev_cold) * 1e-3) / 1e6,
        "gpu_load_ms": gpu_load_ms,
        "gpu_load_steps": load_steps,
        "kernel_ms_isolated": kern_ms,
        "launch_ms": launch_ms,
        "bytes_per_lup": b,
        "state_bytes_per_launch": nodes * b,
        "hbm_frac_whole_job": mlups * 1e6 * b / (HBM_PEAK_GBS * 1e9),
        "kernel_gbs": achieved,
        "kernel_hbm_frac": achieved / HBM_PEAK_GBS,
        "nusselt_hot": d["nusselt_hot"],
        "nusselt_cold": d["nusselt_cold"],
        "mass": d["total_mass"],
        "max_u": d["max_u"],
        "all_finite": d["all_finite"],
        "time_step": d["time_step"],
    }
    s.close()
    return res


def batched_form(steps: int, warmup: int, storage: str = "fp32", total_steps: int = 50000):
    """One launch per step for all cavities (graph replays), autotuned, against
    the same cavities stepped one after another with the single-cavity path.
    After the timed steps the batch runs on to total_steps (the workload's
    specified length) and reports how many cavities stay finite, the largest
    max |u| and the largest relative mass drift."""
    from cfd_cavity_amd import CavityBatch, CavitySolver
    from cfd_cavity_amd.batch import seeded_batch_specs

    specs = seeded_batch_specs()
    b = CavityBatch(specs, storage=storage)
    nodes = b.nodes()
    b.run(warmup)
    tune = b.autotune(trial_steps=max(130, steps // 10))
    ms = b.time_run(steps)
    mlups = nodes * steps / (ms * 1e-3) / 1e6
    finite = sum(1 for c in b if c.diagnostics()["all_finite"])
    run_on = {}
    if total_steps > 0:
        done = b[0].time_step
        if total_steps > done:
            t0 = time.perf_counter()
            b.run(total_steps - done)
            wall = time.perf_counter() - t0
        else:
            wall = 0.0
        diags = [(c.diagnostics(), sp["nx"] * sp["ny"]) for c, sp in zip(b, specs)]
        run_on = {"total_steps": diags[0][0]["time_step"],
                  "finite_cavities": sum(1 for d, _ in diags if d["all_finite"]),
                  "max_u": max(d["max_u"] for d, _ in diags),
                  "max_rel_mass_drift": max(abs(d["total_mass"] - n) / n for d, n in diags),
                  "run_on_wall_s": round(wall, 3)}
    b.close()
    seq_ms = 0.0
    for sp in specs:
        s = CavitySolver(sp["nx"], sp["ny"], sp["ra"], sp["pr"], sp["cs"], sp["pr_t"],
                         sp["utarget"], sp["seed"], 1, storage=storage)
        s.run(warmup)
        seq_ms += s.time_run(steps)
        s.close()
    seq_mlups = nodes * steps / (seq_ms * 1e-3) / 1e6
    return {"cavities": len(specs), "nodes": nodes, "storage": storage, "steps": steps,
            "mlups": mlups, "ms_per_step": ms / steps, "sequential_mlups": seq_mlups,
            "speedup_vs_sequential": mlups / seq_mlups, "tuned": tune,
            "finite_cavities": finite, "run_on": run_on,
            # MLUPS x the one-step algorithm's bytes per update over the HBM
            # peak: an effective rate, which exceeds 1 once several steps
            # run per launch (the state then crosses HBM once per N updates)
            "effective_hbm_frac": mlups * 1e6 * BYTES_PER_LUP[storage] / (HBM_PEAK_GBS * 1e9)}


def measured_traffic(steps_per_launch: int):
    """HBM bytes per FP64 step-kernel launch at 256x1024 (the N-step kernel
    or the single-step one) from the committed rocprofv3 counter summaries
    (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy kernel), or None
    when the summary is absent."""
    name, key = (("k_stepn_256x1024.md", f"TRAFFIC_FP64_STEP{steps_per_launch}_BYTES_PER_LAUNCH")
                 if steps_per_launch > 1 else
                 ("k_step_fused_256x1024.md", "TRAFFIC_FP64_BYTES_PER_LAUNCH"))
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith(key):
                return int(line.split("=")[1].strip())
    except OSError:
        pass
    return None


CLOCK_GHZ = 2.4          # MI355X peak engine clock
# shader clock measured under a dense FP64 FMA load on every CU (s_memtime
# against the 100 MHz s_memrealtime in a probe kernel, two boxes)
CLOCK_GHZ_FP64_LOAD = 2.0
SIMDS = 256 * 4          # 256 CUs x 4 SIMDs
FP64_CYCLES_PER_WAVE_INST = 4   # 16 FP64 lanes per SIMD: a wave64 instruction in 4 cycles


def fp64_issue_bound(steps_per_launch: int, launch_ms: float):
    """The second bound of the multi-step kernel: its FP64 VALU issue time
    (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 per launch, from the committed
    counter summary, at 4 cycles per wave instruction over the 1024 SIMDs at
    the peak clock and at the clock measured under FP64 load) as a fraction
    of the launch time, and the same for all of its VALU instructions
    (SQ_INSTS_VALU); None when the summary holds no figure for this
    kernel."""
    keys = {"fp64": f"FP64_WAVE_INSTS_STEP{steps_per_launch}_PER_LAUNCH",
            "valu": f"VALU_WAVE_INSTS_STEP{steps_per_launch}_PER_LAUNCH"}
    found = {}
    try:
        for line in open(os.path.join(ROOT, "profiles", "k_stepn_256x1024.md")):
            for name, key in keys.items():
                if line.startswith(key):
                    found[name] = float(line.split("=")[1].strip())
    except OSError:
        return None
    if "fp64" not in found:
        return None

    def busy(n, ghz=CLOCK_GHZ):
        return n * FP64_CYCLES_PER_WAVE_INST / SIMDS / (ghz * 1e6)

    n = found["fp64"]
    out = {"wave_instructions_per_launch": n, "issue_ms": round(busy(n), 6),
           "issue_frac": round(busy(n) / launch_ms, 4),
           "clock_ghz": CLOCK_GHZ,
           "issue_frac_at_fp64_load_clock": round(busy(n, CLOCK_GHZ_FP64_LOAD) / launch_ms, 4),
           "fp64_load_clock_ghz": CLOCK_GHZ_FP64_LOAD,
           "source": "profiles/k_stepn_256x1024.md"}
    if "valu" in found:
        v = found["valu"]
        out.update({"all_valu_wave_instructions_per_launch": v,
                    "all_valu_issue_ms": round(busy(v), 6),
                    "all_valu_issue_frac": round(busy(v) / launch_ms, 4),
                    "all_valu_issue_frac_at_fp64_load_clock":
                        round(busy(v, CLOCK_GHZ_FP64_LOAD) / launch_ms, 4)})
    return out


def clock_probe():
    """The 4-step kernel's launch time against the GPU's recent activity, for
    a kernel trace (rocprofv3 --kernel-trace): a fresh solver's first 100
    steps, 8000 steps of sustained load, a second fresh solver (early flow
    state) right after that load, the loaded solver again, and again after
    one second of idle.  Phases are printed as they start."""
    s = make_solver("fp64")
    s2 = make_solver("fp64")
    s2.run(5)
    s2.synchronize()
    phases = [("fresh solver, 100 steps", s, 100), ("sustained load, 8000 steps", s, 8000),
              ("early-state solver right after the load, 100 steps", s2, 100),
              ("loaded solver, 400 steps", s, 400)]
    for i, (name, solver, n) in enumerate(phases, 1):
        print(f"[clock-probe] phase {i}: {name}", file=sys.stderr, flush=True)
        solver.run(n)
        solver.synchronize()
    time.sleep(1.0)
    print("[clock-probe] phase 5: the same solver after 1 s idle, 400 steps", file=sys.stderr,
          flush=True)
    s.run(400)
    s.synchronize()
    s.close()
    s2.close()


def copy_probe(mib: int, reps: int = 5):
    """A plain device copy of known size (torch's copy kernel), for
    calibrating FETCH_SIZE / WRITE_SIZE: run it under rocprofv3 --pmc."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "reps": reps}), flush=True)


def settle(s, ms: float) -> int:
    """Run the solver itself for `ms` of sustained load, which brings the GPU
    to the clocks a long run sees; returns the steps run."""
    t0 = time.perf_counter()
    n = 0
    while (time.perf_counter() - t0) * 1e3 < ms:
        s.run(200)
        s.synchronize()
        n += 200
    return n


def dynamic_form(steps: int, warmup: int, cpu_s: float, gpu_load_ms: float = 300.0):
    """The dynamic mixed-scale SGS path (solver.rs:562-828) on the same
    256x1024 workload, FP64: one launch per step (streaming, the model and
    the collision per tile), against the oracle's dynamic step on a bounded
    CPU sample; and the FP32-storage step in its one-launch and four-launch
    forms.  Each form is timed right after `warmup` steps of a fresh solver
    (*_after_idle, the GPU coming out of idle, the first steps of the
    spin-up), then, for the reported figures, after `gpu_load_ms` of
    sustained load on the same solver (settled clocks, past the spin-up:
    graph replay)."""
    from oracle import pyoracle

    s = make_solver("fp64")
    nodes = WORKLOAD["nx"] * WORKLOAD["ny"]
    s.set_dynamic_sgs(True)
    s.run(warmup)
    ms_idle = s.time_run(steps)
    load_steps = settle(s, gpu_load_ms)
    ms = s.time_run(steps)
    out = {"steps": steps, "ms_per_step": ms / steps, "mlups": nodes * steps / ms / 1e3,
           "ms_per_step_after_idle": ms_idle / steps, "gpu_load_ms": gpu_load_ms,
           "gpu_load_steps": load_steps,
           "launches_per_step": s.dynamic_launches, "cs_dyn_mean": s.dynamic_stats()[0]}
    # FP32 storage: the model, pull and collision in one launch per step, and
    # the four-launch form
    for fuse in (True, False):
        f = make_solver("fp32")
        f.set_dynamic_fusion(fuse)
        f.set_dynamic_sgs(True)
        f.run(warmup)
        ms_idle = f.time_run(steps)
        settle(f, gpu_load_ms)
        ms = f.time_run(steps)
        out[f"fp32_{f.dynamic_launches}_launches_ms_per_step"] = ms / steps
        out[f"fp32_{f.dynamic_launches}_launches_ms_per_step_after_idle"] = ms_idle / steps
        del f
    if cpu_s > 0:
        threads = min(os.cpu_count() or 1, 16)
        w = WORKLOAD
        o = pyoracle.OracleSolver(w["nx"], w["ny"], w["ra"], w["pr"], w["cs"], w["pr_t"],
                                  w["utarget"], w["seed"], threads)
        o.set_dynamic_sgs(True)
        o.run(1)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_s:
            o.run(1)
            n += 1
        dt = time.perf_counter() - t0
        out["cpu_mlups"] = nodes * n / dt / 1e6
        out["cpu_sample"] = f"{n} dynamic steps, {threads} threads (collision), model serial"
        out["speedup_vs_cpu"] = out["mlups"] / out["cpu_mlups"]
    return out


def octree_form(levels: int, steps: int, cpu_s: float, cfg=(64, 256, 8), dynamic=False):
    """The octree cavity (src/octree, the reference CLI's default 64x256 base
    grid with bl_width 8, Ra=1e6) for 2 or 3 levels, optionally with the
    dynamic mixed-scale SGS: coarse steps per second and node updates per
    second (a coarse step updates L0 once, the L1 blocks twice and the L2
    strips four times), against the octree oracle (one thread) on a bounded
    CPU sample."""
    from cfd_cavity_amd import OctreeCavitySolver
    from oracle import pyoracle

    nx0, ny0, bl = cfg
    g = OctreeCavitySolver(nx0, ny0, bl, 1.0e6, levels=levels)
    g.set_dynamic_sgs(dynamic)
    upd = sum(nx * ny * (1, 2, 2, 4, 4)[gi] for gi, (nx, ny) in enumerate(g.dims))
    g.run(20)
    ms = g.time_run(steps)
    out = {"config": f"{nx0}x{ny0}, bl_width {bl}, Ra=1e6, levels {levels}"
                     + (", dynamic SGS" if dynamic else ""),
           "us_per_coarse_step": ms / steps * 1e3, "node_updates_per_step": upd,
           "mnups": upd * steps / ms / 1e3}
    if cpu_s > 0:
        o = pyoracle.OracleOctree(nx0, ny0, bl, 1.0e6, levels=levels)
        o.set_dynamic_sgs(dynamic)
        o.run(1)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < cpu_s:
            o.run(5)
            n += 5
        dt = time.perf_counter() - t0
        out["cpu_mnups"] = upd * n / dt / 1e6
        out["cpu_sample"] = f"{n} coarse steps, oracle/octree_oracle.c, 1 thread"
        out["speedup_vs_cpu"] = out["mnups"] / out["cpu_mnups"]
    return out


def run_loop_form(samples: int = 50, every: int = 48):
    """The run loop's helpers on the FP64 256x1024 workload: one sample of
    the on-device time averages (post.rs:11-57) from the C-form state (the
    M-form fields are materialised first, then k_avg_acc adds u, v, T, nu_t
    and u^2 to five FP64 sum planes), at a cadence of `every` steps; the
    accumulation alone (state already materialised; 112 algorithmic bytes per
    node: 4 planes and 5 sums read, 5 sums written); and a CAVSNAP1 snapshot
    written atomically from the device fields and read back into a restart
    (snapshot.rs:10-131, main.rs:104-128, 284-300).  Host clocks around
    synchronised regions, settled clocks (the solver runs for `gpu_load_ms`
    first)."""
    import tempfile

    import numpy as np

    from cfd_cavity_amd import snapshot

    nodes = WORKLOAD["nx"] * WORKLOAD["ny"]
    s = make_solver("fp64")
    settle(s, 300.0)

    def region(fn):
        s.synchronize()
        t0 = time.perf_counter()
        fn()
        s.synchronize()
        return time.perf_counter() - t0

    def steps_only():
        for _ in range(samples):
            s.run(every)

    def steps_and_samples():
        for _ in range(samples):
            s.run(every)
            s.average_accumulate()

    def accumulate_only():
        for _ in range(4 * samples):
            s.average_accumulate()

    s.average_reset()
    base = min(region(steps_only) for _ in range(3))
    with_samples = min(region(steps_and_samples) for _ in range(3))
    acc = min(region(accumulate_only) for _ in range(3)) / (4 * samples)
    sample = (with_samples - base) / samples
    out = {"config": "256x1024 Ra=1e10 FP64, a sample every %d steps" % every,
           "ms_per_step_without_samples": base * 1e3 / (samples * every),
           "sample_us": sample * 1e6,
           "sample_over_step": sample / (base / (samples * every)),
           "accumulate_us": acc * 1e6,
           "accumulate_gbs": 112.0 * nodes / acc / 1e9,
           "average_count": s.average_count()}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "snap.bin")
        t0 = time.perf_counter()
        snapshot.save(s, path)
        out["snapshot_write_ms"] = (time.perf_counter() - t0) * 1e3
        out["snapshot_bytes"] = os.path.getsize(path)
        r = make_solver("fp64")
        t0 = time.perf_counter()
        snapshot.restart(r, path)
        r.synchronize()
        out["restart_ms"] = (time.perf_counter() - t0) * 1e3
        same = all(np.array_equal(np.asarray(s.get_field(n)), np.asarray(r.get_field(n)))
                   for n in ("rho", "ux", "uy", "temperature"))
        out["restart_fields_equal"] = bool(same)
        r.close()
    s.close()
    return out


def adjoint_form():
    """The adjoint form's workload: 96x384, Ra=1e7, FP64, J = Nu_hot averaged
    over steps 8k-10k from the seeded spun-up state, theta = 0.5 +- 0.01
    (seeded), Cs = 0.10, checkpoints every 100 steps.  Wall time of one
    gradient (dJ/dtheta over 384 nodes and dJ/dCs by the adjoint), of the
    forward window alone, and of one tangent direction."""
    import numpy as np
    from cfd_cavity_amd import CavitySolver
    from cfd_cavity_amd.adjoint import adjoint_nu_window, nu_window, tangent_nu_window

    s = CavitySolver(96, 384, 1e7, 0.71, 0.10, 0.9, 0.02, 42, 1)
    s.run(8000)
    s.synchronize()
    theta = 0.5 + 0.01 * np.random.default_rng(42).uniform(-1, 1, s.ny())

    def timed(f):
        t0 = time.perf_counter()
        r = f()
        return r, time.perf_counter() - t0

    nu_window(s, theta, 0.10, 2000)                       # warm-up (graphs, buffers)
    (j, tf) = timed(lambda: nu_window(s, theta, 0.10, 2000))
    adjoint_nu_window(s, theta, 0.10, 2000, 100)
    (ja, ta) = timed(lambda: adjoint_nu_window(s, theta, 0.10, 2000, 100))
    dth = np.random.default_rng(1).standard_normal(s.ny())
    tangent_nu_window(s, theta, 0.10, dth, 0.0, 2000)
    (_, tt) = timed(lambda: tangent_nu_window(s, theta, 0.10, dth, 0.0, 2000))
    return {"config": "96x384, Ra=1e7, FP64, J = mean Nu_hot over steps 8k-10k, "
                      "checkpoints every 100 steps",
            "forward_window_ms": tf * 1e3, "adjoint_gradient_ms": ta * 1e3,
            "adjoint_over_forward": ta / tf, "tangent_direction_ms": tt * 1e3,
            "gradient_entries": s.ny() + 1, "j": j, "j_adjoint": ja[0]}


def cpu_rate(threads: int, budget_s: float):
    """Steps of the oracle on the 256x1024 workload from the reference init
    until the time budget is spent: (MLUPS, steps, seconds)."""
    from oracle import pyoracle

    w = WORKLOAD
    o = pyoracle.OracleSolver(w["nx"], w["ny"], w["ra"], w["pr"], w["cs"], w["pr_t"],
                              w["utarget"], w["seed"], threads)
    nodes = w["nx"] * w["ny"]
    o.run(2)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.run(5 if threads > 1 else 1)
        steps += 5 if threads > 1 else 1
    dt = time.perf_counter() - t0
    return nodes * steps / dt / 1e6, steps, dt


def cpu_baseline(budget_s: float):
    """The oracle (row-band OpenMP restatement of the rayon path) on a bounded
    sample of the same workload, with all host threads (max 16), and on a
    shorter sample with one thread."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    # 16 threads: the host CPU share of one GPU on the measurement boxes
    threads = min(affinity or os.cpu_count() or 1, 16)
    v, steps, dt = cpu_rate(threads, budget_s)
    v1, steps1, dt1 = cpu_rate(1, max(budget_s / 4, 1.0))
    return {"value": v, "unit": "MLUPS", "cores": threads, "kind": "port",
            "host_cpu_count": os.cpu_count(), "affinity_cpu_count": affinity,
            "sample": f"{steps} steps of 256x1024 Ra=1e10 from the reference init "
                      f"({dt:.1f} s, oracle/cavity_oracle.c, {threads} OpenMP row-band threads)",
            "value_1thread": v1,
            "sample_1thread": f"{steps1} steps, {dt1:.1f} s, 1 thread",
            "reference_published": "14.1 (1 thread) -> 32.0 (8 threads) MLUPS, M1 Max"}


def copy_bandwidth():
    """Device-to-device copies of 80 MB (the FP64 state's size) and 1 GiB,
    timed with HIP events: the practical HBM ceiling on this box (GB/s of
    read + write bytes)."""
    import torch

    out = {}
    for name, nbytes in (("copy_80MB", 79691776), ("copy_1GiB", 1 << 30)):
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        reps = 50 if nbytes < (1 << 28) else 10
        e0.record()
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name + "_gbs"] = 2 * nbytes / (ms * 1e-3) / 1e9
        del a, b
    return out


GATE_CASES = {
    # steady cavities (20k steps) and the 256x1024 Ra=1e10 configuration after
    # 10k steps (plumes formed, before the flow turns chaotic)
    "64x256_ra1e6": (dict(nx=64, ny=256, ra=1.0e6, utarget=0.02), 20000),
    "96x384_ra1e7": (dict(nx=96, ny=384, ra=1.0e7, utarget=0.02), 20000),
    "256x1024_ra1e10": (dict(nx=256, ny=1024, ra=1.0e10, utarget=0.005), 10000),
}


def precision_gate():
    """The metric's "mixed vs FP64 at rel dNu <= 1e-4": per case, FP32 and
    FP16-offset storage against FP64 from the same init after the same steps
    (rel |dNu| over the hot and cold walls, rel |dw_max| at mid-height, mass
    drift)."""
    from cfd_cavity_amd import CavitySolver

    out = {}
    for name, (c, steps) in GATE_CASES.items():
        sols = {st: CavitySolver(c["nx"], c["ny"], c["ra"], 0.71, 0.10, 0.9, c["utarget"], 42, 1,
                                 storage=st) for st in ("fp64", "fp32", "fp16_offset")}
        m0 = c["nx"] * c["ny"]
        vals = {}
        for st, s in sols.items():
            s.run(steps)
            d = s.diagnostics()
            w, _ = s.max_uy_on_row(c["ny"] // 2)
            vals[st] = (d["nusselt_hot"], d["nusselt_cold"], w, abs(d["total_mass"] - m0) / m0)
            s.close()
        r = vals["fp64"]
        row = {"steps": steps, "nusselt_hot_fp64": r[0], "mass_drift_fp64": r[3]}
        for st in ("fp32", "fp16_offset"):
            v = vals[st]
            row[st] = {"rel_dNu": max(abs(v[0] - r[0]) / abs(r[0]), abs(v[1] - r[1]) / abs(r[1])),
                       "rel_dw_max_mid": abs(v[2] - r[2]) / abs(r[2]), "mass_drift": v[3]}
            row[st]["meets_1e-4"] = row[st]["rel_dNu"] <= 1e-4
        out[name] = row
    return out


def precision_run(steps: int, batches: int = 10):
    """The precision comparison at 256x1024, Ra=1e10 over `steps` steps: FP64,
    FP32 and FP16-offset storage from the same init, plus an FP64 run from a
    different seed (44; seeds 42 and 43 give the same state, the generator
    is seeded with seed | 1) as the control of the flow's own variability.
    Nu_hot is sampled every 1000 steps over the second half; its mean comes
    with the standard error from `batches` batch means.  Per form: Nu (hot,
    cold) at the end, the second-half mean and its sigma, w_max on the
    mid-height row ny/2 (src/main.rs:325-330), the relative mass drift; the
    other runs also their differences to the FP64 run, with the difference of
    the means in units of the combined sigma."""
    import numpy as np

    nodes = WORKLOAD["nx"] * WORKLOAD["ny"]
    out = {"steps": steps, "sample_every": 1000, "batches": batches}
    runs = (("fp64", "fp64", 42), ("fp32", "fp32", 42), ("fp16_offset", "fp16_offset", 42),
            ("fp64_seed44", "fp64", 44))
    for name, storage, seed in runs:
        w = WORKLOAD
        from cfd_cavity_amd import CavitySolver

        s = CavitySolver(w["nx"], w["ny"], w["ra"], w["pr"], w["cs"], w["pr_t"], w["utarget"],
                         seed, 1, storage=storage)
        m0 = s.diagnostics()["total_mass"]
        s.run(steps // 2)
        samples = []
        for _ in range(steps // 2 // 1000):
            s.run(1000)
            samples.append(s.diagnostics()["nusselt_hot"])
        s.run(steps - steps // 2 - 1000 * len(samples))
        d = s.diagnostics()
        wmax, _ = s.max_uy_on_row(WORKLOAD["ny"] // 2)
        x = np.array(samples)
        nb = min(batches, max(len(x), 1))
        bm = np.array([v.mean() for v in np.array_split(x, nb)]) if len(x) else np.zeros(1)
        sigma = float(bm.std(ddof=1) / np.sqrt(nb)) if nb > 1 else float("nan")
        out[name] = {"nusselt_hot": d["nusselt_hot"], "nusselt_cold": d["nusselt_cold"],
                     "mean_nusselt_hot_2nd_half": float(x.mean()) if len(x) else float("nan"),
                     "sigma_mean_nusselt_hot": sigma, "w_max_mid": wmax,
                     "mass_drift_rel": abs(d["total_mass"] - m0) / m0,
                     "all_finite": d["all_finite"], "time_step": d["time_step"]}
        s.close()
    ref = out["fp64"]
    for name in ("fp32", "fp16_offset", "fp64_seed44"):
        v = out[name]
        for key in ("nusselt_hot", "mean_nusselt_hot_2nd_half", "w_max_mid"):
            v["rel_err_" + key] = abs(v[key] - ref[key]) / abs(ref[key])
        comb = (v["sigma_mean_nusselt_hot"] ** 2 + ref["sigma_mean_nusselt_hot"] ** 2) ** 0.5
        v["mean_diff_in_sigma"] = (abs(v["mean_nusselt_hot_2nd_half"]
                                       - ref["mean_nusselt_hot_2nd_half"]) / comb
                                   if comb > 0 else float("nan"))
    out["nodes"] = nodes
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--kernel-launches", type=int, default=200)
    ap.add_argument("--forms", default="fp64,fp32,fp16_offset,fp64_2step,fp64_1step")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gpu-load-ms", type=float, default=300.0,
                    help="sustained load on a throwaway solver before each form's timed "
                         "steps (0: none); the form's after-idle figures are measured first")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--batch-total-steps", type=int, default=50000,
                    help="run the batch on to this many steps after timing (0: no)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the dynamic-SGS and octree measurements")
    ap.add_argument("--octree-levels", default="2,3", help="octree configurations to measure")
    ap.add_argument("--octree-forms", default="plain,dynamic",
                    help="octree forms: plain (each level count) and dynamic (2 levels)")
    ap.add_argument("--precision-steps", type=int, default=200000,
                    help="steps of the FP64/FP32/FP16-offset comparison (0: skip)")
    ap.add_argument("--paths", default="uniform,batched,dynamic_sgs,octree,adjoint,run_loop",
                    help="paths to measure (a subset prints a partial line, for profiling)")
    ap.add_argument("--clock-probe", action="store_true",
                    help="only run the launch-time-against-GPU-activity phases (for a trace)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.clock_probe:
        clock_probe()
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {a.gpus} requested); the multi-GPU "
              "layout is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    # torch (device synchronisation, copies) brings its own HIP runtime; it is
    # loaded before the first solver loads the library's, as torch's device
    # initialisation fails when the other runtime was initialised first
    import torch  # noqa: F401

    paths = set(a.paths.split(","))
    if "uniform" not in paths:
        a.forms = ""
    if "batched" not in paths:
        a.no_batch = True
    forms = {}
    for storage in filter(None, a.forms.split(",")):
        print(f"[bench] {storage}: warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
        forms[storage] = gpu_form(storage, a.steps, a.warmup, a.kernel_launches,
                                  gpu_load_ms=a.gpu_load_ms)
    main_form = forms.get("fp64") or next(iter(forms.values()), None)
    ref_nu = forms["fp64"]["nusselt_hot"] if "fp64" in forms else None
    for k, v in forms.items():
        if ref_nu is not None:
            v["rel_dNu_vs_fp64"] = abs(v["nusselt_hot"] - ref_nu) / abs(ref_nu)
            v["speedup_vs_fp64"] = v["mlups"] / forms["fp64"]["mlups"]
            v["rel_dmass_vs_fp64"] = abs(v["mass"] - forms["fp64"]["mass"]) / forms["fp64"]["mass"]

    batched = None
    if not a.no_batch:
        print(f"[bench] batched: 64 cavities, {a.batch_steps} steps", file=sys.stderr, flush=True)
        batched = batched_form(a.batch_steps, a.warmup, total_steps=a.batch_total_steps)

    extra = None
    if not a.no_extra:
        cs = 0.0 if a.no_cpu else 3.0
        extra = {}
        if "dynamic_sgs" in paths:
            print("[bench] dynamic SGS 256x1024", file=sys.stderr, flush=True)
            extra["dynamic_sgs"] = dynamic_form(500, 20, cs, a.gpu_load_ms)
        if "octree" in paths:
            oforms = set(a.octree_forms.split(","))
            for lv in (int(x) for x in a.octree_levels.split(",")):
                if "plain" in oforms:
                    print(f"[bench] octree, {lv} levels", file=sys.stderr, flush=True)
                    extra[f"octree_levels{lv}"] = octree_form(lv, 1000, cs)
            if "dynamic" in oforms:
                print("[bench] octree, 2 levels, dynamic SGS", file=sys.stderr, flush=True)
                extra["octree_levels2_dynamic"] = octree_form(2, 500, cs, dynamic=True)
        if "adjoint" in paths:
            print("[bench] adjoint form 96x384", file=sys.stderr, flush=True)
            extra["adjoint"] = adjoint_form()
        if "run_loop" in paths:
            print("[bench] run loop: time averages and snapshots", file=sys.stderr, flush=True)
            extra["run_loop"] = run_loop_form()

    precision = gate = None
    if "uniform" in paths and a.precision_steps > 0:
        print("[bench] precision gate (FP32 / FP16-offset vs FP64)", file=sys.stderr, flush=True)
        gate = precision_gate()
        print(f"[bench] precision forms, {a.precision_steps} steps", file=sys.stderr, flush=True)
        precision = precision_run(a.precision_steps)
    copy = copy_bandwidth() if main_form is not None else None

    cpu = None
    if not a.no_cpu and main_form is not None:
        print(f"[bench] cpu baseline ({a.cpu_seconds} s)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(a.cpu_seconds)
    if main_form is None:  # a partial run for profiling: no headline figures
        print(json.dumps({"paths": sorted(paths), "batched": batched,
                          "other_paths": extra}), flush=True)
        return

    nodes = WORKLOAD["nx"] * WORKLOAD["ny"]
    b = main_form["bytes_per_lup"]
    spl = main_form["steps_per_launch"]
    traffic = measured_traffic(spl)
    line = {
        "metric": METRIC,
        "value": round(main_form["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_form["ms_per_step"],
        "ms_per_step_wall": main_form["ms_per_step_wall"],
        "value_wall": round(main_form["mlups_wall"], 2),
        "ms_per_step_after_idle": main_form["ms_per_step_after_idle"],
        "value_after_idle": round(main_form["mlups_after_idle"], 2),
        "clock_state": (f"timed right after {a.gpu_load_ms:g} ms of sustained load on a "
                        "throwaway solver of the same workload (clocks settled as in a long "
                        "run); *_after_idle: the same measurement right after set-up"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(main_form["mlups"] / REFERENCE_CPU_MLUPS_256x1024, 2),
        "dtype": DTYPE["fp64" if "fp64" in forms else storage_of(next(iter(forms)))],
        "data": "synthetic (reference init: linear T + seeded xorshift64* noise)",
        "config": {"workload": "thermal LBM D2Q9 f+g DDF cavity 256x1024, Ra=1e10, Pr=0.71, "
                               "Pr_t=0.9, Cs=0.10, utarget=0.005, seed 42; TRT+Smagorinsky+Guo, "
                               "pull streaming, bounce-back, isothermal/adiabatic walls",
                   "nx": WORKLOAD["nx"], "ny": WORKLOAD["ny"], "nodes": nodes,
                   "storage": "fp64"},
        "roofline": {"bound": "hbm", "achieved": round(main_form["kernel_gbs"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(main_form["kernel_hbm_frac"], 4),
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": spl * nodes * b,
                     "traffic_bytes_per_lup": (round(traffic / (spl * nodes), 1)
                                               if traffic else None),
                     "launch_ms": main_form["launch_ms"], "steps_per_launch": spl,
                     "timing": "HIP events on the solver stream around each of 11 repetitions "
                               "of the timed steps; median; settled clocks (clock_state)",
                     "fp64_valu": fp64_issue_bound(spl, main_form["launch_ms"])},
        "cpu_baseline": cpu,
        "forms": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in forms.items()},
        "batched": batched,
        "precision_gate": gate,
        "precision_forms": precision,
        "copy_bandwidth": copy,
        "other_paths": ({k: {kk: (round(vv, 6) if isinstance(vv, float) else vv)
                             for kk, vv in v.items()} for k, v in extra.items()}
                        if extra else None),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
