"""The uniform-mode driver (cfd_cavity_amd/cavity_les, C++ over the C-ABI):
flag validation and exit codes of src/cli.rs:334-383 and src/main.rs (CPU),
and a short GPU run with averaging, snapshots, restart and the abort guard."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cfd_cavity_amd", "cavity_les")


def run(*args, timeout=120):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("args,needle", [
    (["--nx", "4"], "≥ 8"),
    (["--nx", "64", "--ny", "64"], "ny doit être > nx"),
    (["--ra", "0"], "ra et pr"),
    (["--sample-every", "0"], "sample-every"),
    (["--jobs", "0"], "jobs"),
    (["--alpha-mult", "-1"], "alpha-mult"),
    (["--nu-start-mult", "0.5"], "nu-start-mult"),
    (["--nu", "0.01"], "0.002"),
    (["--nx"], "manquante"),
    (["--bogus", "1"], "inconnue"),
    (["--mode", "quadtree"], "mode"),
    (["--mode", "octree", "--nx0", "10"], "nx0 et ny0"),
    (["--mode", "octree", "--nx0", "64", "--ny0", "64"], "ny0 doit être > nx0"),
    (["--mode", "octree", "--bl-width", "40"], "bl-width"),
    (["--mode", "octree", "--levels", "4"], "levels doit valoir"),
    (["--mode", "octree", "--odd-exact=maybe"], "booléen"),
    (["--smag-alpha", "1.5"], "smag-alpha"),
    (["--eta-test", "0.5"], "eta-test"),
    (["--dynamic-smag=maybe"], "booléen"),
])
def test_cli_validation_exit_1(args, needle):
    r = run(*args)
    assert r.returncode == 1 and needle in r.stderr


def test_cli_help():
    r = run("--help")
    assert r.returncode == 0 and "Usage" in r.stdout


@pytest.mark.gpu
def test_cli_run_averaging_snapshot_and_log(tmp_path):
    from cfd_cavity_amd import CavitySolver, snapshot

    out = tmp_path / "run"
    r = run("--nx", 32, "--ny", 128, "--ra", 1e6, "--steps", 2500, "--sample-start", 1000,
            "--sample-every", 100, "--progress-every", 1000, "--snapshot-every", 1000,
            "--out", out)
    assert r.returncode == 0, r.stderr
    lines = (out / "progress.log").read_text().splitlines()
    assert [ln.split("|")[0].strip() for ln in lines] == ["step       0", "step    1000",
                                                         "step    2000", "step    2499"]
    summ = json.loads((out / "summary.json").read_text())
    assert summ["time_samples"] == 15 and summ["steps"] == 2500
    nx, ny, rho, ux, uy, t = snapshot.read_snapshot(out / "snapshot.bin")
    # cavity_les's final state equals the library stepped the same number of steps
    s = CavitySolver(32, 128, 1e6, 0.71, 0.1, 0.9, 0.02, 42, 1)
    s.run(2500)
    assert (nx, ny) == (32, 128) and np.array_equal(t, s.get_field("temperature"))
    d = s.diagnostics()
    assert f"Nuc = {d['nusselt_hot']:.4f}" in lines[-1]
    # post.rs:202-316 exports: mean fields with the stream function, the w
    # profiles, the final fields as VTK ImageData
    fm = np.loadtxt(out / "fields_mean.csv", delimiter=",", skiprows=1)
    assert fm.shape == (nx * ny, 7)
    assert np.array_equal(fm[:, 0], np.tile(np.arange(nx), ny))
    ux_m = fm[:, 2].reshape(ny, nx)
    psi = fm[:, 5].reshape(ny, nx)
    ref_psi = np.vstack([np.zeros((1, nx)), np.cumsum(0.5 * (ux_m[1:] + ux_m[:-1]), axis=0)])
    # the file holds 9 significant digits of ux and psi: the column sums of
    # the rounded ux agree to ~1e-8 of the largest |psi|
    assert np.max(np.abs(psi - ref_psi)) < 1e-6 * np.max(np.abs(ref_psi))
    pw = np.loadtxt(out / "profiles_w.csv", delimiter=",", skiprows=1)
    assert pw.shape == (nx, 3) and pw[0, 0] == 0.0 and pw[-1, 0] == 1.0
    assert np.allclose(pw[:, 1], fm[:, 3].reshape(ny, nx)[ny // 2], rtol=1e-7, atol=1e-14)
    vti = (out / "snapshot.vti").read_text()
    assert vti.startswith('<?xml version="1.0"?>') and f'WholeExtent="0 {nx - 1} 0 {ny - 1} 0 0"' in vti
    tv = np.array(vti.split('Name="Temperature" format="ascii">')[1].split("</DataArray>")[0]
                  .split(), dtype=float)
    assert np.allclose(tv, s.get_field("temperature"), rtol=1e-6, atol=1e-12)


@pytest.mark.gpu
def test_cli_restart_errors_and_abort(tmp_path):
    from cfd_cavity_amd import snapshot

    r = run("--restart", tmp_path / "missing.bin", "--steps", 10, "--out", tmp_path / "o")
    assert r.returncode == 6
    nx, ny = 32, 128
    n = nx * ny
    ux = np.full(n, 0.2)
    t = np.array([0.5 - (k % nx) / (nx - 1) for k in range(n)])
    path = tmp_path / "fast.bin"
    snapshot.write_snapshot(path, nx, ny, np.ones(n), ux, np.zeros(n), t)
    r = run("--nx", nx, "--ny", ny, "--restart", path, "--steps", 100, "--out", tmp_path / "a")
    assert r.returncode == 2 and "AVORT" in r.stderr
    assert "AVORT" in (tmp_path / "a" / "progress.log").read_text()


@pytest.mark.gpu
def test_cli_dynamic_smag(tmp_path):
    """--dynamic-smag (main.rs:94-102, :179-187, :393-424): the announcement,
    the coefficient suffix of the progress lines and dynamic_sgs.json; a
    following flag is not taken as the switch's value (cli.rs:201-211)."""
    out = tmp_path / "dyn"
    r = run("--nx", 24, "--ny", 96, "--ra", 1e6, "--steps", 600, "--sample-start", 300,
            "--sample-every", 50, "--progress-every", 200, "--dynamic-smag", "--eta-test", 3,
            "--out", out)
    assert r.returncode == 0, r.stderr
    assert "LES dynamique : échelles mixtes α = 0.5 (Tab. 2.3) | η_test = 3 | νsm_min = -0.1·ν" \
        in r.stdout
    lines = (out / "progress.log").read_text().splitlines()
    assert all(" | cs_dyn_mean = " in ln and " | kap_dyn_mean = " in ln for ln in lines)
    dj = json.loads((out / "dynamic_sgs.json").read_text())
    assert dj["dynamic_smag"] is True and dj["eta_test"] == 3 and dj["smag_alpha"] == 0.5
    assert 0.0 <= dj["cs_dyn_max"] <= 0.04 and dj["cs_dyn_mean"] <= dj["cs_dyn_max"]
    assert "LES dynamique : cs_dyn_mean = " in r.stdout


@pytest.mark.gpu
def test_cli_octree_run_snapshot_and_restart(tmp_path):
    """--mode octree (main.rs:445-738): progress lines, OCTSNAP1 exports, the
    V2 summary, and a restart from the exported snapshot."""
    from cfd_cavity_amd import OctreeCavitySolver
    from cfd_cavity_amd.snapshot import read_octree_snapshot

    out = tmp_path / "oct"
    args = ["--mode", "octree", "--nx0", 32, "--ny0", 64, "--bl-width", 4, "--levels", 3,
            "--sample-start", 100, "--sample-every", 50, "--progress-every", 100,
            "--snapshot-every", 100]
    r = run(*args, "--steps", 300, "--out", out)
    assert r.returncode == 0, r.stderr
    lines = (out / "progress.log").read_text().splitlines()
    assert [ln.split("|")[0].strip() for ln in lines[:4]] == ["step       0", "step     100",
                                                             "step     200", "step     299"]
    assert "max_rho_dev_ifaces = " in lines[0] and "Nu_fin = " in lines[0]
    assert "PORTE V2" in r.stdout and "--- Bilan (moyenné sur 4 échantillons" in r.stdout
    levels, step, grids = read_octree_snapshot(out / "snapshot.bin")
    assert (levels, step, len(grids)) == (3, 300, 5)
    s = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    s.run(300)
    for gi, g in enumerate(grids):
        assert np.array_equal(g[5], s.get_field(gi, "temperature"))
    r2 = run(*args, "--steps", 50, "--sample-start", 0, "--restart", out / "snapshot.bin",
             "--out", tmp_path / "oct2")
    assert r2.returncode == 0, r2.stderr
    assert "Restart depuis" in r2.stdout and "pas 300" in r2.stdout
    bad = run("--mode", "octree", "--nx0", 32, "--ny0", 64, "--bl-width", 6, "--steps", 5,
              "--restart", out / "snapshot.bin", "--out", tmp_path / "oct3")
    assert bad.returncode == 6 and "incompatible" in bad.stderr


@pytest.mark.gpu
def test_cli_octree_dynamic_smag(tmp_path):
    """--mode octree --dynamic-smag (main.rs:515-523, :627-641): the
    announcement and the cs_dyn / kap_dyn / nu_t_max suffix of every progress
    line, with the values of the library's octree stepped alike."""
    from cfd_cavity_amd import OctreeCavitySolver

    out = tmp_path / "octdyn"
    r = run("--mode", "octree", "--nx0", 32, "--ny0", 64, "--bl-width", 4, "--levels", 3,
            "--steps", 120, "--sample-start", 0, "--sample-every", 40, "--progress-every", 40,
            "--dynamic-smag", "--out", out)
    assert r.returncode == 0, r.stderr
    assert "LES dynamique (Étape D) : alpha_mixte = 0.5 | nu_sm_min = -0.1 | eta_test = 2" \
        in r.stdout
    lines = (out / "progress.log").read_text().splitlines()[:4]
    assert all(" | cs_dyn = " in ln and " | kap_dyn = " in ln and " | nu_t_max = " in ln
               for ln in lines)
    s = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    s.set_dynamic_sgs(True)
    s.run(120)
    cs, _, kap, _, _ = s.dynamic_stats()
    assert f"cs_dyn = {cs:.4f} | kap_dyn = {kap:.4f}" in lines[-1]


@pytest.mark.gpu
@pytest.mark.parametrize("storage,tol", [("fp32", 1e-3), ("fp16_offset", 2e-2)])
def test_cli_dynamic_smag_low_precision(tmp_path, storage, tol):
    """--storage fp32 / fp16_offset with --dynamic-smag (the one-launch step in
    that storage, the model in FP64): the run completes with the coefficient
    suffix on every progress line, and its time-averaged Nusselt number stays
    within the stated relative tolerance of the FP64 run."""
    nus = {}
    for st in ("fp64", storage):
        out = tmp_path / st
        r = run("--nx", 24, "--ny", 96, "--ra", 1e6, "--steps", 1200, "--sample-start", 600,
                "--sample-every", 50, "--progress-every", 400, "--dynamic-smag",
                "--storage", st, "--out", out)
        assert r.returncode == 0, r.stderr
        lines = (out / "progress.log").read_text().splitlines()
        assert lines and all(" | cs_dyn_mean = " in ln for ln in lines)
        nus[st] = json.loads((out / "summary.json").read_text())["nusselt_hot"]
    assert abs(nus[storage] - nus["fp64"]) <= tol * abs(nus["fp64"]), nus
