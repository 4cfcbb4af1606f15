"""Shared test helpers: error metrics and canonical octree forms (order-independent)."""
from __future__ import annotations

import hashlib

import numpy as np

CANON_DTYPE = np.dtype([("lo", "<i8"), ("level", "<i8"), ("hi", "<i8"), ("nchild", "<i8"), ("leaf", "<i8")])


def rel_rms(a, ref) -> float:
    """sqrt(sum |a - ref|^2 / sum |ref|^2), the reference tests' metric (tests/test_cpu_backend.cpp:62-67)."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    den = float((ref ** 2).sum())
    num = float(((a - ref) ** 2).sum())
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def per_body_rel(a, ref) -> np.ndarray:
    a = np.asarray(a, np.float64).reshape(-1, 3)
    ref = np.asarray(ref, np.float64).reshape(-1, 3)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _canon(lo, level, hi, nchild, leaf) -> np.ndarray:
    out = np.zeros(len(lo), CANON_DTYPE)
    out["lo"], out["level"], out["hi"], out["nchild"], out["leaf"] = lo, level, hi, nchild, leaf
    out.sort(order=["lo", "level"])
    return out


def canon_from_ref_export(t: dict) -> np.ndarray:
    """Canonical cells from the reference's breadth-first whole-tree export (LetNode records,
    src/backend.hpp:63-87): body ranges follow from leaf body counts, children in octant order."""
    k = len(t["child_count"])
    fc, cc, bc = t["first_child"].astype(np.int64), t["child_count"].astype(np.int64), t["body_count"].astype(np.int64)
    count = np.where(fc < 0, bc, 0).astype(np.int64)
    for i in range(k - 1, -1, -1):
        if fc[i] >= 0:
            count[i] = count[fc[i]:fc[i] + cc[i]].sum()
    lo = np.zeros(k, np.int64)
    level = np.zeros(k, np.int64)
    for i in range(k):
        if fc[i] >= 0:
            o = lo[i]
            for c in range(fc[i], fc[i] + cc[i]):
                lo[c] = o
                level[c] = level[i] + 1
                o += count[c]
    leaf = (fc < 0).astype(np.int64)
    return _canon(lo, level, lo + count, np.where(fc < 0, 0, cc), leaf)


def canon_from_oracle(nd: dict) -> np.ndarray:
    nchild = nd["nchild"].astype(np.int64)
    return _canon(nd["lo"], nd["depth"], nd["hi"], nchild, (nchild == 0).astype(np.int64))


def canon_from_hip(nd: dict) -> np.ndarray:
    return _canon(nd["lo"], nd["level"], nd["hi"], nd["nchild"], nd["leaf"])


def morton_keys_np(x, y, z):
    """Closed-form restatement of the key recipe (src/cpu_backend.cpp:404-426) in numpy float64:
    origin = float min, side = max extent * (1 + 1e-6) (1 if 0), q = clamp((v - o) * 2^21/side)."""
    x, y, z = (np.asarray(v, np.float32) for v in (x, y, z))
    lo = [float(v.min()) for v in (x, y, z)]
    hi = [float(v.max()) for v in (x, y, z)]
    side = max(float(np.float64(hi[c]) - np.float64(lo[c])) for c in range(3))
    side = side * (1 + 1e-6) if side > 0 else 1.0
    s = float(2 ** 21) / side

    def q(v, o):
        t = (v.astype(np.float64) - o) * s
        return np.clip(t, 0.0, float(2 ** 21 - 1)).astype(np.uint64)

    def dil(v):
        r = np.zeros_like(v)
        for b in range(21):
            r |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return r

    return (dil(q(x, lo[0])) << np.uint64(2)) | (dil(q(y, lo[1])) << np.uint64(1)) | dil(q(z, lo[2]))


def restart_positions(kind: str, seed: int = 1) -> np.ndarray:
    """3000 float32 positions in [-1, 1]^3 with finest-level cells holding more than 16 bodies that do
    not coincide, so the build restarts keys in those cells' own cubes (src/cpu_backend.cpp:501-544):
    "clump": 200 bodies within 1e-6 of a point (one restart frame);
    "nested": a body at 1000 sets a large root cube; 399 bodies in [0, 2e-3]^3, 50 within 2e-9 of a
    point and 9 copies at each of 21 consecutive float x values (restarts inside a restarted cell);
    "coincident_in_restart": 100 bodies within 1e-6 of a point, 40 more on one of them (a coincident
    leaf inside a restarted cell)."""
    rng = np.random.default_rng(seed)
    n = 3000
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    if kind == "clump":
        x[:200] = (np.float32([0.3, -0.2, 0.1]) + rng.uniform(0, 1e-6, (200, 3))).astype(np.float32)
    elif kind == "nested":
        x[0] = [1000.0, 1000.0, 1000.0]
        x[1:400] = rng.uniform(0, 2e-3, (399, 3)).astype(np.float32)
        x[400:450] = (np.float32(1e-3) + rng.uniform(0, 2e-9, (50, 3))).astype(np.float32)
        v = np.float32(1.5e-3)
        for k in range(21):
            x[450 + 9 * k: 459 + 9 * k] = [v, np.float32(1.5e-3), np.float32(1.5e-3)]
            v = np.nextafter(v, np.float32(1))
    elif kind == "coincident_in_restart":
        x[:100] = (np.float32([0.5, 0.5, 0.5]) + rng.uniform(0, 1e-6, (100, 3))).astype(np.float32)
        x[100:140] = x[0]
    else:
        raise ValueError(kind)
    return x
