"""GPU tests of snapshot restarts (src/simulation.cpp:176-215, 311-325): a run snapshotted mid-way
and restarted in a fresh context continues bit for bit, in the reference's f32 state and, through
the FP64 extension of the file, in the mixed-precision form."""
from __future__ import annotations

import numpy as np
import pytest

import lettuce_simulate_physics_amd as L

pytestmark = pytest.mark.gpu

DT = 1e-3


def step(be, p, k, potential=False):
    """k KDK steps as the reference driver runs them (src/simulation.cpp:406-420)."""
    for _ in range(k):
        be.kick(0.5 * DT)
        be.drift(DT)
        be.compute_forces(p, potential)
        be.kick(0.5 * DT)


def state(be, mixed):
    """Positions and velocities by id (FP64 in the mixed form)."""
    if mixed:
        x, v, _, ids = be.download_f64()
    else:
        b = be.download()
        x, v, ids = np.stack([b.x, b.y, b.z], 1), np.stack([b.vx, b.vy, b.vz], 1), b.id
    o = np.argsort(ids)
    return x[o], v[o]


@pytest.fixture(scope="module")
def second():
    be, why = L.make_hip_backend(0)
    if be is None:
        pytest.fail(f"HIP backend unavailable on a GPU test run: {why}")
    yield be
    be.close()


@pytest.mark.parametrize("mode", ["plain", "mixed"])
@pytest.mark.parametrize("method", ["tree", "direct"])
def test_restart_continues_bit_for_bit(hip, second, tmp_path, mode, method):
    mixed = mode == "mixed"
    n = 20000 if method == "tree" else 4096
    p = L.ForceParams(G=1.0, eps=0.01, theta=0.6, method=method)
    path = str(tmp_path / f"snap_{mode}_{method}.snap")
    for be in (hip, second):
        be.set_precision(mode)
    try:
        if mixed:
            x, v, m = L.plummer_f64(n, 7)
            hip.upload_f64(x, v, m)
        else:
            hip.upload(L.plummer(n, 7))
        hip.compute_forces(p, False)
        step(hip, p, 6)
        hip.write_snapshot(path, step=6, time=6 * DT, G=1.0, eps=0.01, dt=DT, seed=7, scenario="plummer")
        step(hip, p, 6)
        want = state(hip, mixed)
        h = second.restore_snapshot(path)
        assert h["n"] == n and h["step"] == 6 and h["fp64"] == mixed
        second.compute_forces(p, False)  # as at step 6 of the first run
        step(second, p, 6)
        got = state(second, mixed)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    finally:
        for be in (hip, second):
            be.set_precision("plain")


def test_fp64_snapshot_in_a_plain_context_and_f32_in_a_mixed_one(hip, second, tmp_path):
    x, v, m = L.plummer_f64(5000, 9)
    x = x + 100.0
    path64, path32 = str(tmp_path / "s64.snap"), str(tmp_path / "s32.snap")
    hip.set_precision("mixed")
    try:
        hip.upload_f64(x, v, m)
        hip.write_snapshot(path64, step=1)
        xr, vr, _, ids = hip.download_f64()
        assert L.snapshot_header(path64)["fp64"]
        second.set_precision("plain")
        second.restore_snapshot(path64)  # a plain context takes the rounded f32 arrays
        b = second.download()
        assert np.array_equal(b.id, ids) and np.array_equal(b.x, xr[:, 0].astype(np.float32))
        second.write_snapshot(path32, step=1)
        assert not L.snapshot_header(path32)["fp64"]
        second.set_precision("mixed")
        second.restore_snapshot(path32)  # an f32 file in a mixed context: promoted
        x2, _, _, _ = second.download_f64()
        assert np.array_equal(x2[:, 0], b.x.astype(np.float64))
    finally:
        hip.set_precision("plain")
        second.set_precision("plain")
