"""CAVSNAP1 snapshot format and restart regrid (src/uniform/snapshot.rs),
restating the reference's snapshot tests (src/uniform/tests.rs:18-111) on the
library's host-side I/O (no GPU)."""
import os

import numpy as np
import pytest

from cfd_cavity_amd import snapshot
from cfd_cavity_amd._lib import CavlbmError, E_IO


def fields(nx, ny, a=1e-4):
    n = nx * ny
    k = np.arange(n, dtype=np.float64)
    return (1.0 + a * k, 0.001 * k, -0.002 * k, 0.5 - k / n)


def test_snapshot_atomic_write_roundtrip_and_no_leftovers(tmp_path):   # tests.rs:18-57
    nx, ny = 10, 14
    rho, ux, uy, t = fields(nx, ny)
    path = tmp_path / "snapshot.bin"
    snapshot.write_snapshot_atomic(path, nx, ny, rho, ux, uy, t)
    rnx, rny, rrho, rux, ruy, rt = snapshot.read_snapshot(path)
    assert (rnx, rny) == (nx, ny)
    for a, b in ((rrho, rho), (rux, ux), (ruy, uy), (rt, t)):
        assert np.array_equal(a, b)
    rho2 = np.full(nx * ny, 2.0)
    snapshot.write_snapshot_atomic(path, nx, ny, rho2, ux, uy, t)
    assert np.array_equal(snapshot.read_snapshot(path)[2], rho2)
    assert [p for p in os.listdir(tmp_path) if p.endswith(".tmp")] == []


def test_snapshot_roundtrip_and_layout(tmp_path):                      # tests.rs:59-82
    nx, ny = 12, 20
    rho, ux, uy, t = fields(nx, ny, 1e-3)
    path = tmp_path / "s.bin"
    snapshot.write_snapshot(path, nx, ny, rho, ux, uy, t)
    raw = path.read_bytes()
    assert raw[:8] == b"CAVSNAP1" and len(raw) == 24 + 32 * nx * ny
    assert int.from_bytes(raw[8:16], "little") == nx and int.from_bytes(raw[16:24], "little") == ny
    assert np.array_equal(np.frombuffer(raw[24:24 + 8 * nx * ny], dtype="<f8"), rho)
    got = snapshot.read_snapshot(path)
    assert np.array_equal(got[5], t)


def test_snapshot_rejects_bad_files(tmp_path):                        # snapshot.rs:62-78
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NOTASNAP" + bytes(16))
    with pytest.raises(CavlbmError) as e:
        snapshot.read_snapshot(bad)
    assert e.value.code == E_IO
    good = tmp_path / "good.bin"
    snapshot.write_snapshot(good, 4, 5, *fields(4, 5))
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(good.read_bytes()[:-8])
    with pytest.raises(CavlbmError):
        snapshot.read_snapshot(trunc)


def test_interpolate_same_grid_identity():                            # tests.rs:84-89
    src = np.ones(5 * 9)
    assert np.array_equal(snapshot.interpolate_field(src, 5, 9, 5, 9), src)


def test_interpolate_linear_field_exact():                            # tests.rs:91-111
    nx0, ny0 = 8, 32
    t0 = np.array([0.5 + (-1.0) * (ix / (nx0 - 1)) for iy in range(ny0) for ix in range(nx0)])
    t1 = snapshot.interpolate_field(t0, nx0, ny0, 16, 64).reshape(64, 16)
    exact = 0.5 - np.arange(16) / 15.0
    assert np.max(np.abs(t1 - exact[None, :])) < 1e-10
    assert np.all(np.abs(t1[:, 0] - 0.5) < 1e-10) and np.all(np.abs(t1[:, -1] + 0.5) < 1e-10)


def test_interpolate_bilinear_in_both_directions():
    """f(x, y) = a + b x + c y + d x y is reproduced exactly by bilinear
    interpolation on any grid change (normalised coordinates)."""
    nx0, ny0, nx1, ny1 = 7, 11, 13, 5
    x0, y0 = np.meshgrid(np.arange(nx0) / (nx0 - 1), np.arange(ny0) / (ny0 - 1))
    f0 = 0.3 + 0.7 * x0 - 0.2 * y0 + 0.9 * x0 * y0
    x1, y1 = np.meshgrid(np.arange(nx1) / (nx1 - 1), np.arange(ny1) / (ny1 - 1))
    f1 = snapshot.interpolate_field(f0.ravel(), nx0, ny0, nx1, ny1).reshape(ny1, nx1)
    assert np.max(np.abs(f1 - (0.3 + 0.7 * x1 - 0.2 * y1 + 0.9 * x1 * y1))) < 1e-12


@pytest.mark.parametrize("nx,ny", [(0, 16), (16, 0), (1 << 32, 1 << 32), ((1 << 62) + 1, 4),
                                   (1 << 60, 1)])
def test_snapshot_corrupt_header_is_rejected(tmp_path, nx, ny):
    """A header whose grid is empty, or whose nx*ny or 32*nx*ny bytes do not
    fit in 64 bits (nx = ny = 2^32 wraps nx*ny to 0), is an I/O error, not an
    empty or truncated read."""
    import struct

    path = tmp_path / "bad.bin"
    path.write_bytes(b"CAVSNAP1" + struct.pack("<QQ", nx, ny) + b"\0" * 64)
    with pytest.raises(CavlbmError) as e:
        snapshot.read_snapshot(path)
    assert e.value.code == E_IO
