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
