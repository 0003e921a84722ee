"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/cavlbm.h declares, its struct layouts match the header, and argument
validation that precedes any device work behaves like the reference's
asserts.  No compute calls (no GPU in this test)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from cfd_cavity_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cavlbm.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(cavlbm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"missing export {n}"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signature table out of sync with header"


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "cavlbm.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(cavlbm_params),"
        " sizeof(cavlbm_diag), sizeof(cavlbm_derived), offsetof(cavlbm_params, nu_override),"
        " offsetof(cavlbm_diag, time_step));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(_lib.Params)
    assert int(out[1]) == C.sizeof(_lib.Diag)
    assert int(out[2]) == C.sizeof(_lib.Derived)
    assert int(out[3]) == _lib.Params.nu_override.offset
    assert int(out[4]) == _lib.Diag.time_step.offset


def test_abi_version_and_defaults():
    L = _lib.lib()
    assert L.cavlbm_abi_version() == 1
    p = _lib.Params()
    assert L.cavlbm_params_default(C.byref(p)) == 0
    # src/cli.rs:94-130
    assert (p.nx, p.ny, p.ra, p.pr, p.cs, p.pr_t, p.utarget, p.seed, p.jobs) == (
        64, 256, 1.0e6, 0.71, 0.10, 0.9, 0.02, 42, 1)
    assert p.alpha_mult == 1.0 and p.nu_start_mult == 1.0 and p.has_nu_override == 0
    assert p.geq_order == 2 and p.storage == 0


def _create(**kw):
    L = _lib.lib()
    p = _lib.Params()
    L.cavlbm_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    h = C.c_void_p()
    rc = L.cavlbm_create(C.byref(p), C.byref(h))
    return rc, L.cavlbm_last_error(None).decode(), h


@pytest.mark.parametrize("kw,needle", [
    (dict(nx=2), "nx and ny"),
    (dict(storage=7), "storage"),
    (dict(geq_order=3), "geq_order"),
    (dict(has_nu_override=1, nu_override=-1.0), "--nu"),
    (dict(has_nu_override=1, nu_override=1e-5), "tau+"),
    (dict(pr=1e4), "tau_t="),                                  # solver.rs:154
])
def test_create_rejects_like_reference_asserts(kw, needle):
    rc, msg, h = _create(**kw)
    assert rc == _lib.E_PARAM and needle in msg and not h.value


def _zero_args(args):
    out = []
    for a in args:
        if a is C.c_double:
            out.append(0.0)
        elif a in (C.c_int32, C.c_uint32, C.c_uint64):
            out.append(0)
        else:
            out.append(None)
    return out


def test_every_handle_entry_point_rejects_a_null_handle():
    """A null solver / batch / octree handle gives CAVLBM_E_PARAM before any
    device work (destroy of null is a no-op returning 0, like free)."""
    L = _lib.lib()
    checked = 0
    for name, (res, args) in _lib.SIGNATURES.items():
        if res is not C.c_int or not args or args[0] is not C.c_void_p:
            continue
        rc = getattr(L, name)(*_zero_args(args))
        want = 0 if name.endswith("_destroy") else _lib.E_PARAM
        assert rc == want, name
        checked += 1
    assert checked >= 50
    for name in ("cavlbm_last_error", "cavlbm_batch_last_error", "cavlbm_octree_last_error"):
        assert isinstance(getattr(L, name)(None), bytes)


def test_create_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    rc, msg, h = _create()
    assert rc in (_lib.E_NODEVICE, _lib.E_HIP) and not h.value and msg


def test_no_cpu_fallback_in_product():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "cfd_cavity_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower().replace("the oracle is", ""), f


def test_missing_library_fails_loudly():
    """Without the built library the package raises on first use: there is no
    CPU fallback for the product path."""
    import sys

    code = ("import cfd_cavity_amd._lib as L; L.LIB_PATH = '/nonexistent/libcavlbm.so'\n"
            "try:\n    L.lib()\nexcept RuntimeError as e:\n    print('raised:', e)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=root, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "raised:" in p.stdout and "no CPU fallback" in p.stdout
