"""CPU tests of the product's boundary: the C-ABI library loads, exports every entry point the
header declares, and behaves as the reference factory does without a device."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import lettuce_simulate_physics_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lhip_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("lhip_create", "lhip_upload", "lhip_download", "lhip_compute_forces", "lhip_kick",
                     "lhip_drift", "lhip_diagnostics", "lhip_exact_potential_energy", "lhip_times",
                     "lhip_local_bounds", "lhip_build_tree", "lhip_compute_forces_with", "lhip_memory_need",
                     "lhip_device_memory", "lhip_body_costs"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = L.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_library_is_built_for_gfx950():
    data = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_issue_probe_library_is_built_for_gfx950():
    # The issue-rate probe bench.py runs beside the force path: its own library, entry issue_probe_run.
    path = os.path.join(ROOT, "lettuce_simulate_physics_amd", "libissue_probe.so")
    lib = ctypes.CDLL(path)
    assert hasattr(lib, "issue_probe_run") and hasattr(lib, "issue_probe_sweep")
    assert b"gfx950" in open(path, "rb").read()
    # The sweep refuses forms and chain counts it has no kernel for before any device call.
    lib.issue_probe_sweep.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double(-1.0)
    assert lib.issue_probe_sweep(5, 12, 1024, 16, ctypes.byref(ms)) == -1
    assert lib.issue_probe_sweep(0, 7, 1024, 16, ctypes.byref(ms)) == -1
    assert lib.issue_probe_sweep(0, 12, 0, 16, ctypes.byref(ms)) == -1
    assert ms.value == -1.0


def test_memory_need_is_host_only_and_monotone():
    d1, h1 = L.HipBackend.memory_need(1 << 20, "tree")
    d2, _ = L.HipBackend.memory_need(1 << 22, "tree")
    dd, _ = L.HipBackend.memory_need(1 << 20, "direct")
    assert d2 > d1 > 0 and dd > 0 and h1 == 0


def test_memory_need_per_body_in_each_mode():
    # Per-body slopes of the reserves: 104 B of f32 state and posm; the tree adds 44 B per body (keys,
    # levels, scans, groups, walk order) and 132 B per node of the first node reserve n / 4
    # (a tree with more nodes grows it); the mixed form adds 128 B of FP64 state (tree: posd 32 B and
    # 80 B per node of FP64 moments; direct: 48-B records).
    # Direct from 2^21 on: below it the plain sum splits its source range and holds the partial sums.
    n2 = 1 << 22
    slope = {}
    for m in ("tree", "direct"):
        n1 = 1 << 20 if m == "tree" else 1 << 21
        for mode in ("plain", "mixed"):
            a, _ = L.HipBackend.memory_need(n1, m, mode)
            b, _ = L.HipBackend.memory_need(n2, m, mode)
            slope[m, mode] = (b - a) / (n2 - n1)
    assert 181 < slope["tree", "plain"] < 182  # the reference's CUDA backend: ~112 B (src/cuda_backend.cu:8-18)
    assert 361 < slope["tree", "mixed"] < 362
    assert 103.5 < slope["direct", "plain"] < 104.5
    assert 311 < slope["direct", "mixed"] < 313
    with pytest.raises(ValueError):
        L.HipBackend.memory_need(n2, "tree", "fp16")


def test_factory_without_device_returns_reason(monkeypatch):
    if L.hip_device_count() > 0:
        pytest.skip("a device is present")
    be, why = L.make_hip_backend(0)
    assert be is None and why


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="missing"):
        L.library()


def test_force_params_validation():
    with pytest.raises(ValueError):
        L.ForceParams(method="auto").c()
    c = L.ForceParams(G=2.0, eps=0.1, theta=0.5, method="tree").c()
    assert c.method == L.METHOD_TREE and c.G == 2.0
    assert ctypes.sizeof(c) == 32
