"""lettuce HIP backend for MI355X (gfx950): Python mirror of the reference ``Backend`` interface.

The compute path is ``liblettuce_hip.so`` (hand-written HIP kernels behind the C ABI declared in
``include/lettuce_hip.h``). This module binds that ABI with ctypes and mirrors the reference's
``Backend`` members (src/backend.hpp:89-141) and factory (``make_cuda_backend``,
src/cuda_backend.cu:2026-2048) with the same names, argument meaning and error behaviour: device
and runtime failures raise ``RuntimeError``, protocol misuse raises ``LogicError``. There is no CPU
fallback: without the library, or without a device, ``make_hip_backend`` returns ``(None, why)``
and ``HipBackend(...)`` raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblettuce_hip.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_U64 = ctypes.c_uint64

OK, RUNTIME, LOGIC, UNAVAILABLE, OUT_OF_MEMORY, INVALID, NOT_PROVIDED = range(7)
METHOD_DIRECT, METHOD_TREE = 0, 1
MODE_PLAIN, MODE_MIXED = 0, 1
REFERENCE_GROUP_MAX = 128  # src/cpu_backend.cpp:27


class LogicError(RuntimeError):
    """Protocol misuse (the reference throws std::logic_error)."""


class ForceParamsC(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("eps", ctypes.c_double), ("theta", ctypes.c_double),
                ("method", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DiagC(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("ke", ctypes.c_double), ("pe", ctypes.c_double),
                ("has_pe", ctypes.c_int32), ("pad", ctypes.c_int32)] + \
               [(k, ctypes.c_double) for k in ("px", "py", "pz", "lx", "ly", "lz", "cx", "cy", "cz")]


class TimesC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("force_ms", "sort_ms", "build_ms", "walk_ms", "step_ms")] + \
               [("interactions", ctypes.c_uint64), ("comm_ms", ctypes.c_double), ("balance_ms", ctypes.c_double),
                ("p2p", ctypes.c_uint64), ("p2c", ctypes.c_uint64), ("listed", ctypes.c_uint64)]


_lib = None


def library():
    """The loaded HIP library. Raises if it has not been built (no silent fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with lettuce_simulate_physics_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        L.lhip_create.argtypes = [ctypes.c_int, ctypes.POINTER(_P), ctypes.c_char_p, ctypes.c_size_t]
        L.lhip_destroy.argtypes = [_P]
        L.lhip_last_error.restype = ctypes.c_char_p
        L.lhip_last_error.argtypes = [_P]
        L.lhip_size.restype = _U64
        L.lhip_size.argtypes = [_P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


@dataclass
class Bodies:
    """Host SoA bodies as the reference's Bodies (src/bodies.hpp:8-13): float x..m, uint32 id."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    m: np.ndarray
    id: np.ndarray

    @staticmethod
    def empty(n: int) -> "Bodies":
        return Bodies(*[np.zeros(n, np.float32) for _ in range(7)], np.zeros(n, np.uint32))

    @staticmethod
    def of(obj) -> "Bodies":
        """From any object with x, y, z, vx, vy, vz, m, id arrays."""
        f = [np.ascontiguousarray(getattr(obj, k), dtype=np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "m")]
        return Bodies(*f, np.ascontiguousarray(obj.id, dtype=np.uint32))

    def arrays(self):
        return [self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m, self.id]

    def __len__(self):
        return len(self.x)


@dataclass
class ForceParams:
    """ForceParams (src/backend.hpp:14-19)."""
    G: float = 1.0
    eps: float = 0.01
    theta: float = 0.6
    method: str = "direct"

    def c(self) -> ForceParamsC:
        if self.method not in ("direct", "tree"):
            raise ValueError(f"method must be 'direct' or 'tree', not {self.method!r}")
        return ForceParamsC(self.G, self.eps, self.theta, METHOD_TREE if self.method == "tree" else METHOD_DIRECT, 0)


def hip_compiled() -> bool:
    try:
        return bool(library().lhip_compiled())
    except RuntimeError:
        return False


def hip_device_count() -> int:
    try:
        return int(library().lhip_device_count())
    except RuntimeError:
        return 0


def make_hip_backend(device: int = 0):
    """Factory with the reference's contract: (backend, "") or (None, why). Never raises."""
    try:
        return HipBackend(device), ""
    except Exception as e:  # noqa: BLE001 - the factory reports every failure as `why`
        return None, str(e)


class HipBackend:
    """The Backend interface over liblettuce_hip.so."""

    def __init__(self, device: int = 0):
        L = library()
        h = _P()
        why = ctypes.create_string_buffer(512)
        rc = L.lhip_create(int(device), ctypes.byref(h), why, len(why))
        if rc != OK:
            raise RuntimeError(why.value.decode() or f"lhip_create failed ({rc})")
        self._L = L
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.lhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc == OK:
            return
        msg = self._L.lhip_last_error(self._h).decode()
        if rc == LOGIC:
            raise LogicError(msg)
        if rc == INVALID:
            raise ValueError(msg or "invalid argument")
        raise RuntimeError(msg or f"lhip error {rc}")

    # ---- Backend members ----
    def name(self) -> str:
        buf = ctypes.create_string_buffer(512)
        self._check(self._L.lhip_name(self._h, buf, len(buf)))
        return buf.value.decode()

    def size(self) -> int:
        return int(self._L.lhip_size(self._h))

    def upload(self, b) -> None:
        b = Bodies.of(b)
        self._check(self._L.lhip_upload(self._h, _U64(len(b)), *[_ptr(a) for a in b.arrays()]))

    def download(self) -> Bodies:
        b = Bodies.empty(self.size())
        self._check(self._L.lhip_download(self._h, *[_ptr(a) for a in b.arrays()]))
        return b

    def compute_forces(self, p: ForceParams, potential: bool = False) -> None:
        c = p.c()
        self._check(self._L.lhip_compute_forces(self._h, ctypes.byref(c), int(potential)))

    def kick(self, dt: float) -> None:
        self._check(self._L.lhip_kick(self._h, _D(dt)))

    def drift(self, dt: float) -> None:
        self._check(self._L.lhip_drift(self._h, _D(dt)))

    def diagnostics(self) -> dict:
        d = DiagC()
        self._check(self._L.lhip_diagnostics(self._h, ctypes.byref(d)))
        out = {k: getattr(d, k) for k, _ in DiagC._fields_ if k != "pad"}
        out["has_pe"] = bool(d.has_pe)
        return out

    def exact_potential_energy(self, G: float, eps: float) -> float:
        v = ctypes.c_double()
        self._check(self._L.lhip_exact_potential_energy(self._h, _D(G), _D(eps), ctypes.byref(v)))
        return v.value

    def times(self) -> dict:
        t = TimesC()
        self._check(self._L.lhip_times(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in TimesC._fields_}

    def local_bounds(self):
        lo = (ctypes.c_float * 3)()
        hi = (ctypes.c_float * 3)()
        self._check(self._L.lhip_local_bounds(self._h, lo, hi))
        return list(lo), list(hi)

    def build_tree(self, p: ForceParams) -> None:
        c = p.c()
        self._check(self._L.lhip_build_tree(self._h, ctypes.byref(c)))

    def compute_forces_with(self, p: ForceParams, potential: bool = False, remote=()) -> None:
        if remote:
            raise NotImplementedError("remote locally essential trees (multi-rank) are not supported on one GPU")
        c = p.c()
        self._check(self._L.lhip_compute_forces_with(self._h, ctypes.byref(c), int(potential)))

    @staticmethod
    def memory_need(n: int, method: str):
        d, h = _U64(), _U64()
        rc = library().lhip_memory_need(_U64(n), METHOD_TREE if method == "tree" else METHOD_DIRECT,
                                        ctypes.byref(d), ctypes.byref(h))
        if rc != OK:
            raise ValueError("memory_need failed")
        return int(d.value), int(h.value)

    def device_memory(self):
        f, t = _U64(), _U64()
        self._check(self._L.lhip_device_memory(self._h, ctypes.byref(f), ctypes.byref(t)))
        return int(f.value), int(t.value)

    def body_costs(self):
        c = np.zeros(self.size(), np.float32)
        rc = self._L.lhip_body_costs(self._h, _ptr(c))
        if rc == NOT_PROVIDED:
            return None
        self._check(rc)
        return c

    # ---- mixed-precision form (no reference counterpart) ----
    def set_precision(self, mode: str = "plain", group_max: int = REFERENCE_GROUP_MAX) -> None:
        """'plain': the reference's FP32 state; 'mixed': FP64 state with FP32 relative-coordinate
        pair arithmetic. group_max: walk group size of later trees (16..128)."""
        if mode not in ("plain", "mixed"):
            raise ValueError("mode must be 'plain' or 'mixed'")
        self._check(self._L.lhip_set_precision(self._h, MODE_MIXED if mode == "mixed" else MODE_PLAIN,
                                               int(group_max)))

    def precision(self):
        m, g = ctypes.c_int32(), ctypes.c_int32()
        self._check(self._L.lhip_get_precision(self._h, ctypes.byref(m), ctypes.byref(g)))
        return ("mixed" if m.value == MODE_MIXED else "plain"), int(g.value)

    def upload_f64(self, x, v, m, ids=None) -> None:
        """FP64 positions and velocities (n x 3), float masses, uint32 ids (default 0..n-1)."""
        n = len(m)
        x = np.asarray(x, np.float64).reshape(n, 3)
        v = np.asarray(v, np.float64).reshape(n, 3)
        cols = [np.ascontiguousarray(a[:, c]) for a in (x, v) for c in range(3)]
        mm = np.ascontiguousarray(m, np.float32)
        ids = np.arange(n, dtype=np.uint32) if ids is None else np.ascontiguousarray(ids, np.uint32)
        self._check(self._L.lhip_upload_f64(self._h, _U64(n), *[_ptr(c) for c in cols], _ptr(mm), _ptr(ids)))

    def download_f64(self):
        """(x (n x 3), v (n x 3), m, id) in the backend's order, FP64 positions and velocities."""
        n = self.size()
        cols = [np.zeros(n) for _ in range(6)]
        mm = np.zeros(n, np.float32)
        ids = np.zeros(n, np.uint32)
        self._check(self._L.lhip_download_f64(self._h, *[_ptr(c) for c in cols], _ptr(mm), _ptr(ids)))
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1), mm, ids

    def accelerations_f64(self):
        n = self.size()
        out = [np.zeros(n) for _ in range(4)]
        self._check(self._L.lhip_accelerations_f64(self._h, *[_ptr(o) for o in out]))
        return out

    # ---- inspection ----
    def accelerations(self):
        n = self.size()
        out = [np.zeros(n, np.float32) for _ in range(4)]
        self._check(self._L.lhip_accelerations(self._h, *[_ptr(o) for o in out]))
        return out

    def tree_keys(self) -> np.ndarray:
        k = np.zeros(self.size(), np.uint64)
        self._check(self._L.lhip_tree_keys(self._h, _ptr(k)))
        return k

    def tree_counts(self):
        a, b, c = _U64(), _U64(), _U64()
        self._check(self._L.lhip_tree_counts(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def tree_nodes(self) -> dict:
        M, _, _ = self.tree_counts()
        link = np.zeros((M, 4), np.int32)
        prefix = np.zeros(M, np.uint64)
        mom = np.zeros((M, 16), np.float32)
        self._check(self._L.lhip_tree_nodes(self._h, _ptr(link), _ptr(prefix), _ptr(mom)))
        return dict(lo=link[:, 0], hi=link[:, 1], level=link[:, 2] & 31, leaf=(link[:, 2] >> 5) & 1,
                    nchild=(link[:, 2] >> 8) & 15, first=link[:, 3], prefix=prefix, com=mom[:, 0:3],
                    m=mom[:, 3], q=mom[:, 4:10], rcrit2=mom[:, 10], geo=mom[:, 12:16])

    def tree_groups(self) -> np.ndarray:
        _, G, _ = self.tree_counts()
        g = np.zeros(G + 1, np.uint32)
        self._check(self._L.lhip_tree_groups(self._h, _ptr(g)))
        return g

    def accel_by_id(self, b, p: ForceParams, potential: bool = False) -> np.ndarray:
        """Accelerations indexed by id, read as the reference tests do: upload with v = 0,
        compute_forces, kick(1), download (tests/test_cuda_backend.cpp:93-112)."""
        z = Bodies.of(b)
        z = Bodies(z.x, z.y, z.z, np.zeros_like(z.vx), np.zeros_like(z.vy), np.zeros_like(z.vz), z.m, z.id)
        self.upload(z)
        self.compute_forces(p, potential)
        self.kick(1.0)
        o = self.download()
        a = np.zeros((len(o), 3))
        a[o.id, 0] = o.vx
        a[o.id, 1] = o.vy
        a[o.id, 2] = o.vz
        return a
