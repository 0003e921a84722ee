"""OCTSNAP1 multi-level snapshots (src/octree/snapshot.rs) through libcavlbm's
host-side I/O (no GPU): the byte layout parsed independently, round trips,
the atomic write and the reference's decode errors."""
import os
import struct

import numpy as np
import pytest

from cfd_cavity_amd._lib import CavlbmError, E_IO
from cfd_cavity_amd.snapshot import read_octree_snapshot, write_octree_snapshot

DIMS3 = [(32, 64), (11, 129), (11, 129), (7, 255), (7, 255)]


def grids(dims, seed=0):
    rng = np.random.default_rng(seed)
    return [(nx, ny) + tuple(rng.standard_normal(nx * ny) for _ in range(4)) for nx, ny in dims]


def test_layout_parsed_independently(tmp_path):          # snapshot.rs:5-17, :57-79
    g = grids(DIMS3[:3], 1)
    p = tmp_path / "s.bin"
    write_octree_snapshot(p, 2, 1234, g)
    b = p.read_bytes()
    assert b[:8] == b"OCTSNAP1"
    version, levels, step, ng = struct.unpack_from("<IIQI", b, 8)
    assert (version, levels, step, ng) == (1, 2, 1234, 3)
    dims = [struct.unpack_from("<QQ", b, 28 + 16 * i) for i in range(ng)]
    assert dims == DIMS3[:3]
    off = 28 + 16 * ng
    for nx, ny, *fields in g:
        for f in fields:
            got = np.frombuffer(b, "<f8", nx * ny, off)
            assert np.array_equal(got, f)
            off += 8 * nx * ny
    assert off == len(b)


def test_roundtrip_levels3_bit_identical(tmp_path):      # tests.rs:795-830
    g = grids(DIMS3, 2)
    a, c = tmp_path / "a.bin", tmp_path / "c.bin"
    write_octree_snapshot(a, 3, 3, g)
    levels, step, g2 = read_octree_snapshot(a)
    assert (levels, step) == (3, 3) and [x[:2] for x in g2] == DIMS3
    for x, y in zip(g, g2):
        for q in range(2, 6):
            assert np.array_equal(x[q], y[q])
    write_octree_snapshot(c, levels, step, g2)
    assert a.read_bytes() == c.read_bytes()
    assert not os.path.exists(str(a) + ".tmp")


def expect_io_error(path, needle):
    with pytest.raises(CavlbmError) as e:
        read_octree_snapshot(path)
    assert e.value.code == E_IO and needle in str(e.value)


def test_decode_errors(tmp_path):                        # snapshot.rs:83-137
    p = tmp_path / "s.bin"
    write_octree_snapshot(p, 2, 7, grids(DIMS3[:3], 3))
    good = p.read_bytes()
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"CAVSNAP1" + good[8:])
    expect_io_error(bad, "magic")
    bad.write_bytes(good[:20])
    expect_io_error(bad, "magic")
    bad.write_bytes(good[:8] + struct.pack("<I", 2) + good[12:])
    expect_io_error(bad, "version 2")
    bad.write_bytes(good[:12] + struct.pack("<I", 3) + good[16:])
    expect_io_error(bad, "levels=3 / n_grids=3")
    bad.write_bytes(good[:40])
    expect_io_error(bad, "tronqué")
    bad.write_bytes(good[:-8])
    expect_io_error(bad, "attendu")
    bad.write_bytes(good + b"\0")
    expect_io_error(bad, "attendu")
    expect_io_error(tmp_path / "missing.bin", "cannot open")
