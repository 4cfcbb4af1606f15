"""Snapshot files (host code of the library, no device needed): the reference's snapshot tests
(tests/test_snapshot.cpp) restated against lhip_snapshot_save/load, byte-level and read-back
compatibility with the reference's own writer and reader (oracle/_ref), and the FP64 extension."""
from __future__ import annotations

import os

import numpy as np
import pytest

import lettuce_simulate_physics_amd as L
import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")

HEADER = dict(step=987654321012, time=12.3456789012345, G=6.5, eps=0.0123, dt=1.0 / 128, seed=5,
              scenario="plummer")


def plummer_bodies(n=1234, seed=5):
    b = L.plummer(n, seed)
    b.id[:] = np.arange(n - 1, -1, -1, dtype=np.uint32)  # ids are data, not positions
    return b


def same_bodies(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a.arrays(), b.arrays()))


def test_bit_exact_round_trip(tmp_path):
    b = plummer_bodies()
    path = str(tmp_path / "snap_rt.bin")
    L.snapshot_save(path, HEADER, b)
    assert not os.path.exists(path + ".tmp")
    assert os.path.getsize(path) >= 4096 * 8
    h, r, f64 = L.snapshot_load(path)
    assert f64 is None and not h["fp64"]
    assert h["n"] == len(b) and all(h[k] == v for k, v in HEADER.items())
    assert same_bodies(r, b)


def test_empty_and_long_scenario_name(tmp_path):
    path = str(tmp_path / "snap_empty.bin")
    L.snapshot_save(path, dict(scenario="q" * 50), L.Bodies.empty(0))
    h, r, _ = L.snapshot_load(path)
    assert h["n"] == 0 and len(r) == 0
    assert h["scenario"] == "q" * 31


def test_rejects_bad_files(tmp_path):
    path = str(tmp_path / "snap_bad.bin")
    L.snapshot_save(path, {}, plummer_bodies(50))
    good = open(path, "rb").read()

    def rejected(data: bytes) -> bool:
        with open(path, "wb") as f:
            f.write(data)
        try:
            L.snapshot_load(path)
        except RuntimeError:
            return True
        return False

    assert rejected(b"X" + good[1:])                       # magic
    assert rejected(good[:8] + b"\x09" + good[9:])         # version
    assert rejected(good[:-8])                             # truncated arrays
    assert rejected(good[:100])                            # truncated header
    assert rejected(b"")
    assert rejected(good[:16] + b"\xff" * 8 + good[24:])  # body count beyond the file
    with pytest.raises(RuntimeError):
        L.snapshot_load(str(tmp_path / "does_not_exist.bin"))


def test_piece_count_and_header_only_read(tmp_path):
    b = plummer_bodies(300)
    path = str(tmp_path / "snap_hdr.bin")
    L.snapshot_save(path, dict(step=40, pieces=3), b)
    h = L.snapshot_header(path)
    assert h["n"] == 300 and h["step"] == 40 and h["pieces"] == 3
    assert L.snapshot_load(path)[0]["pieces"] == 3
    L.snapshot_save(path, dict(step=40), b)  # older files have zeros where the count lives
    assert L.snapshot_header(path)["pieces"] == 0
    os.truncate(path, 4096)  # the header read checks the body count against the file size
    with pytest.raises(RuntimeError):
        L.snapshot_header(path)


# ---------------------------------------------------------------- FP64 extension (no reference counterpart)

def fp64_state(n=777, seed=3):
    x, v, m = L.plummer_f64(n, seed)
    x = x + 1e3  # far from the origin: the float rounding drops many bits
    b = L.Bodies(*[np.ascontiguousarray(a, np.float32) for a in (x[:, 0], x[:, 1], x[:, 2], v[:, 0], v[:, 1],
                                                                  v[:, 2], m)], np.arange(n, dtype=np.uint32))
    return b, x, v


def test_fp64_round_trip_and_rounded_f32(tmp_path):
    b, x, v = fp64_state()
    path = str(tmp_path / "snap64.bin")
    L.snapshot_save(path, HEADER, b, (x, v))
    h, r, f64 = L.snapshot_load(path)
    assert h["fp64"] and f64 is not None
    assert np.array_equal(f64[0], x) and np.array_equal(f64[1], v)  # bit for bit
    assert same_bodies(r, b)
    assert np.array_equal(r.x, x[:, 0].astype(np.float32))            # f32 = round to nearest
    assert (np.float64(r.x) != x[:, 0]).any()                          # the FP64 state is more


def test_fp64_extension_rejects_bad_files(tmp_path):
    b, x, v = fp64_state(100)
    path = str(tmp_path / "snap64_bad.bin")
    L.snapshot_save(path, {}, b, (x, v))
    good = open(path, "rb").read()
    with open(path, "wb") as f:
        f.write(good[:-8])  # the last FP64 array cut short
    with pytest.raises(RuntimeError):
        L.snapshot_load(path)
    with open(path, "wb") as f:
        f.write(good[:172] + b"\x07\x00\x00\x00" + good[176:])  # an extension this reader does not know
    with pytest.raises(RuntimeError):
        L.snapshot_load(path)
    with pytest.raises(ValueError):
        L.snapshot_save(path, {}, b, (x, v[:50]))


# ---------------------------------------------------------------- the reference's writer and reader

@needs_ref
def test_f32_file_is_byte_identical_to_the_reference_writer(tmp_path):
    b = plummer_bodies(5000)
    ours, theirs = str(tmp_path / "ours.snap"), str(tmp_path / "ref.snap")
    hdr = dict(HEADER, pieces=2)
    L.snapshot_save(ours, hdr, b)
    O.ref_write_snapshot(theirs, hdr, O.Bodies(*b.arrays()))
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    h, r, _ = L.snapshot_load(theirs)
    assert same_bodies(r, b) and h["pieces"] == 2 and h["scenario"] == "plummer"


@needs_ref
def test_reference_reads_the_fp64_file(tmp_path):
    b, x, v = fp64_state()
    path = str(tmp_path / "snap64.snap")
    L.snapshot_save(path, dict(HEADER, pieces=1), b, (x, v))
    h, r = O.ref_read_snapshot(path)
    assert h["n"] == len(b) and h["step"] == HEADER["step"] and h["time"] == HEADER["time"] and h["pieces"] == 1
    assert all(np.array_equal(p, q) for p, q in zip(r.arrays(), b.arrays()))
