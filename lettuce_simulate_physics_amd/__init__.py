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
# Walk options (group_max, small_leaf, quad_tau) of each mode when none are given: plain is the
# reference walk (src/cpu_backend.cpp:27,29, every quadrupole kept).
DEFAULT_WALK = {"plain": (REFERENCE_GROUP_MAX, 4, 0.0), "mixed": (REFERENCE_GROUP_MAX, 1, 0.05)}


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
                ("p2p", ctypes.c_uint64), ("p2c", ctypes.c_uint64), ("listed", ctypes.c_uint64),
                ("p2c_mono", ctypes.c_uint64)]


class WalkOptionsC(ctypes.Structure):
    _fields_ = [("group_max", ctypes.c_int32), ("small_leaf", ctypes.c_int32), ("quad_tau", ctypes.c_double)]


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
        L.lhip_snapshot_error.restype = ctypes.c_char_p
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


class DiskParamsC(ctypes.Structure):
    _fields_ = [("nstars", ctypes.c_uint64), ("seed", ctypes.c_uint64)] + \
               [(k, ctypes.c_double) for k in ("G", "core", "mass", "rd", "rmax", "z0", "sigma", "incl", "spin")] + \
               [("pos", ctypes.c_double * 3), ("vel", ctypes.c_double * 3)]


def _check_gen(rc):
    if rc != OK:
        raise ValueError("invalid scenario parameters")


def plummer(n: int, seed: int = 1, a: float = 1.0, cut: float = 10.0, M: float = 1.0, G: float = 1.0) -> Bodies:
    """The reference's Plummer scenario (src/scenarios.cpp:94-132), float bodies bit-exact with it
    (keys plummer.a, plummer.cut, plummer.mass, seed)."""
    b = Bodies.empty(n)
    _check_gen(library().lhip_scenario_plummer(_U64(n), _U64(seed), _D(G), _D(M), _D(a), _D(cut),
                                               *[_ptr(v) for v in b.arrays()]))
    return b


def plummer_f64(n: int, seed: int = 1, a: float = 1.0, cut: float = 10.0, M: float = 1.0, G: float = 1.0):
    """The same sphere before rounding to float: (x (n x 3), v (n x 3), m) in FP64."""
    c = [np.zeros(n) for _ in range(7)]
    _check_gen(library().lhip_scenario_plummer_f64(_U64(n), _U64(seed), _D(G), _D(M), _D(a), _D(cut),
                                                   *[_ptr(v) for v in c], None))
    return np.stack(c[:3], 1), np.stack(c[3:6], 1), c[6]


def _disk_params(nstars, seed, core, mass, rd, rmax, z0, sigma, incl, spin, pos, vel, G):
    return DiskParamsC(nstars, seed, G, core, mass, rd, rmax, z0, sigma, incl, spin, (ctypes.c_double * 3)(*pos),
                       (ctypes.c_double * 3)(*vel))


def disk_sech2(nstars: int, seed: int = 1, core: float = 10.0, mass: float = 1.0, rd: float = 1.0,
               rmax: float = 10.0, z0: float = 0.1, sigma: float = 0.1, incl: float = 0.0, spin: float = 1.0,
               pos=(0.0, 0.0, 0.0), vel=(0.0, 0.0, 0.0), com_frame: bool = True, G: float = 1.0) -> Bodies:
    """Rotating exponential disk with a sech^2 vertical profile around a central point mass (body 0;
    nstars + 1 bodies). No reference counterpart (the reference disk is Gaussian in z)."""
    p = _disk_params(nstars, seed, core, mass, rd, rmax, z0, sigma, incl, spin, pos, vel, G)
    b = Bodies.empty(nstars + 1)
    _check_gen(library().lhip_scenario_disk_sech2(ctypes.byref(p), int(com_frame), *[_ptr(v) for v in b.arrays()]))
    return b


def disk_sech2_f64(nstars: int, seed: int = 1, core: float = 10.0, mass: float = 1.0, rd: float = 1.0,
                   rmax: float = 10.0, z0: float = 0.1, sigma: float = 0.1, incl: float = 0.0, spin: float = 1.0,
                   pos=(0.0, 0.0, 0.0), vel=(0.0, 0.0, 0.0), com_frame: bool = True, G: float = 1.0):
    """The disk in FP64: (x (n x 3), v (n x 3), m)."""
    p = _disk_params(nstars, seed, core, mass, rd, rmax, z0, sigma, incl, spin, pos, vel, G)
    c = [np.zeros(nstars + 1) for _ in range(7)]
    _check_gen(library().lhip_scenario_disk_sech2_f64(ctypes.byref(p), int(com_frame), *[_ptr(v) for v in c], None))
    return np.stack(c[:3], 1), np.stack(c[3:6], 1), c[6]


class CollisionParamsC(ctypes.Structure):
    _fields_ = [("nstars", ctypes.c_uint64), ("seed", ctypes.c_uint64)] + \
               [(k, ctypes.c_double) for k in ("G", "core", "mass", "rd", "rmax", "z0", "sigma", "pericentre",
                                               "separation", "incl1", "incl2")]


@dataclass
class Collision:
    """Two sech^2 disks on a parabolic encounter (see lhip_scenario_collision); the
    differentiable parameters are pericentre, incl1 and incl2 (radians)."""
    nstars: int = 524288
    seed: int = 1
    pericentre: float = 4.0
    incl1: float = np.pi / 6
    incl2: float = np.pi / 3
    separation: float = 20.0
    core: float = 10.0
    mass: float = 1.0
    rd: float = 1.0
    rmax: float = 10.0
    z0: float = 0.1
    sigma: float = 0.1
    G: float = 1.0

    def c(self) -> CollisionParamsC:
        return CollisionParamsC(self.nstars, self.seed, self.G, self.core, self.mass, self.rd, self.rmax, self.z0,
                                self.sigma, self.pericentre, self.separation, self.incl1, self.incl2)

    @property
    def n(self) -> int:
        return 2 * (self.nstars + 1)

    def generate(self):
        """(x (n x 3), v (n x 3), m) in FP64."""
        x, v, m = np.zeros((self.n, 3)), np.zeros((self.n, 3)), np.zeros(self.n)
        p = self.c()
        _check_gen(library().lhip_scenario_collision(ctypes.byref(p), _ptr(x), _ptr(v), _ptr(m), None))
        return x, v, m

    def vjp(self, gx, gv):
        """(dL/dpericentre, dL/dincl1, dL/dincl2) from dL/dx0 and dL/dv0."""
        gx = np.ascontiguousarray(gx, np.float64).reshape(self.n, 3)
        gv = np.ascontiguousarray(gv, np.float64).reshape(self.n, 3)
        out = np.zeros(3)
        p = self.c()
        _check_gen(library().lhip_scenario_collision_vjp(ctypes.byref(p), _ptr(gx), _ptr(gv), _ptr(out)))
        return out


# LetNode (src/backend.hpp:63-76): the 60-byte wire record of a locally essential tree.
LET_NODE = np.dtype([(k, "<f4") for k in ("cx", "cy", "cz", "mass", "qxx", "qxy", "qxz", "qyy", "qyz", "qzz",
                                          "rcrit2")] +
                    [(k, "<i4") for k in ("first_child", "child_count", "body_first", "body_count")])
assert LET_NODE.itemsize == 60


@dataclass
class LetTree:
    """LetTree (src/backend.hpp:84-87): LET_NODE records and the exported bodies (float SoA)."""
    nodes: np.ndarray
    bx: np.ndarray
    by: np.ndarray
    bz: np.ndarray
    bm: np.ndarray

    @staticmethod
    def empty() -> "LetTree":
        return LetTree(np.zeros(0, LET_NODE), *[np.zeros(0, np.float32) for _ in range(4)])

    def __len__(self):
        return len(self.nodes)


class LetC(ctypes.Structure):
    _fields_ = [("nodes", _P), ("n_nodes", ctypes.c_uint64), ("bx", _P), ("by", _P), ("bz", _P), ("bm", _P),
                ("n_bodies", ctypes.c_uint64)]


class SnapshotHeaderC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("step", ctypes.c_uint64), ("time", ctypes.c_double), ("G", ctypes.c_double),
                ("eps", ctypes.c_double), ("dt", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("scenario", ctypes.c_char * 32), ("pieces", ctypes.c_uint32), ("fp64", ctypes.c_uint32)]


SNAPSHOT_FIELDS = ("step", "time", "G", "eps", "dt", "seed", "scenario", "pieces")


def _snap_header(h: dict | None) -> SnapshotHeaderC:
    h = dict(h or {})
    unknown = set(h) - set(SNAPSHOT_FIELDS) - {"n", "fp64"}
    if unknown:
        raise ValueError(f"unknown snapshot header fields: {sorted(unknown)}")
    c = SnapshotHeaderC()
    c.step, c.time, c.G = int(h.get("step", 0)), float(h.get("time", 0.0)), float(h.get("G", 1.0))
    c.eps, c.dt, c.seed = float(h.get("eps", 0.0)), float(h.get("dt", 0.0)), int(h.get("seed", 0))
    c.scenario = str(h.get("scenario", "")).encode()[:31]
    c.pieces = int(h.get("pieces", 0))
    return c


def _snap_dict(c: SnapshotHeaderC) -> dict:
    return dict(n=c.n, step=c.step, time=c.time, G=c.G, eps=c.eps, dt=c.dt, seed=c.seed,
                scenario=c.scenario.decode(), pieces=c.pieces, fp64=bool(c.fp64))


def _snap_check(rc):
    if rc == OK:
        return
    msg = library().lhip_snapshot_error().decode()
    if rc == INVALID:
        raise ValueError(msg or "invalid snapshot argument")
    raise RuntimeError(msg or f"snapshot error {rc}")


def snapshot_header(path: str) -> dict:
    """read_snapshot_header (src/snapshot.hpp:26): the header, validated against the file size."""
    c = SnapshotHeaderC()
    _snap_check(library().lhip_snapshot_read_header(os.fsencode(path), ctypes.byref(c)))
    return _snap_dict(c)


def snapshot_save(path: str, header: dict | None, b, f64=None) -> None:
    """write_snapshot (src/snapshot.hpp:23): bodies b; f64 = (x (n x 3), v (n x 3)) adds the FP64
    state (the f32 arrays should then be its rounding, as download() gives)."""
    b = Bodies.of(b)
    cols = [None] * 6
    if f64 is not None:
        x, v = (np.asarray(a, np.float64).reshape(len(b), 3) for a in f64)
        cols = [np.ascontiguousarray(a[:, c]) for a in (x, v) for c in range(3)]
    c = _snap_header(header)
    _snap_check(library().lhip_snapshot_save(os.fsencode(path), ctypes.byref(c), _U64(len(b)),
                                             *[_ptr(a) for a in b.arrays()], *[_ptr(a) for a in cols]))


def snapshot_load(path: str):
    """read_snapshot (src/snapshot.hpp:24): (header, Bodies, (x, v) in FP64 or None)."""
    n = snapshot_header(path)["n"]
    b = Bodies.empty(n)
    cols = [np.zeros(n) for _ in range(6)]
    c = SnapshotHeaderC()
    _snap_check(library().lhip_snapshot_load(os.fsencode(path), ctypes.byref(c), _U64(n),
                                             *[_ptr(a) for a in b.arrays()], *[_ptr(a) for a in cols]))
    h = _snap_dict(c)
    f64 = (np.stack(cols[:3], 1), np.stack(cols[3:], 1)) if h["fp64"] else None
    return h, b, f64


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


def direct_forces_noalloc(posm: int, n: int, G: float, eps: float, potential: bool, ax: int, ay: int, az: int,
                          pot: int | None = None, stream: int | None = None) -> None:
    """lhip_direct_forces_noalloc: the context's direct sum on caller-owned device arrays (integer
    device addresses: posm holds n float4 (x, y, z, m), ax/ay/az/pot n floats), enqueued on `stream`
    (None: the null stream) without allocating or synchronising. Raises ValueError for arguments the
    library refuses, RuntimeError for launch failures."""
    rc = library().lhip_direct_forces_noalloc(_P(posm), _U64(n), _D(G), _D(eps), 1 if potential else 0, _P(ax),
                                               _P(ay), _P(az), _P(pot) if pot else None, _P(stream) if stream else None)
    if rc == INVALID:
        raise ValueError("lhip_direct_forces_noalloc: n must be below 2^31 - 1, arrays non-null (pot with the "
                         "potential), posm 16-byte aligned")
    if rc != OK:
        raise RuntimeError(f"lhip_direct_forces_noalloc failed with status {rc}")


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
        """Forces from the local bodies plus every remote LetTree (src/backend.hpp:131-136)."""
        c = p.c()
        if not remote:
            self._check(self._L.lhip_compute_forces_with(self._h, ctypes.byref(c), int(potential)))
            return
        keep, lets = [], (LetC * len(remote))()
        for k, t in enumerate(remote):
            nodes = np.ascontiguousarray(t.nodes, LET_NODE)
            cols = [np.ascontiguousarray(v, np.float32) for v in (t.bx, t.by, t.bz, t.bm)]
            if len({len(v) for v in cols}) != 1:
                raise ValueError("LET body arrays differ in length")
            keep += [nodes, *cols]
            lets[k] = LetC(_ptr(nodes), len(nodes), *[_ptr(v) for v in cols], len(cols[0]))
        self._check(self._L.lhip_compute_forces_remote(self._h, ctypes.byref(c), int(potential), lets, len(remote)))

    def export_let(self, box, p: ForceParams) -> LetTree:
        """The part of the local tree a target anywhere in box = (lo, hi) needs (src/backend.hpp:125-130)."""
        lo = (ctypes.c_float * 3)(*[float(v) for v in box[0]])
        hi = (ctypes.c_float * 3)(*[float(v) for v in box[1]])
        c = p.c()
        nn, nb = _U64(), _U64()
        self._check(self._L.lhip_export_let(self._h, lo, hi, ctypes.byref(c), ctypes.byref(nn), ctypes.byref(nb)))
        t = LetTree(np.zeros(nn.value, LET_NODE), *[np.zeros(nb.value, np.float32) for _ in range(4)])
        self._check(self._L.lhip_export_let_copy(self._h, _ptr(t.nodes), _ptr(t.bx), _ptr(t.by), _ptr(t.bz),
                                                 _ptr(t.bm)))
        return t

    @staticmethod
    def memory_need(n: int, method: str, mode: str = "plain"):
        """(device, host) bytes a force call on n bodies reserves (src/backend.hpp:137)."""
        if mode not in ("plain", "mixed"):
            raise ValueError("mode must be 'plain' or 'mixed'")
        d, h = _U64(), _U64()
        rc = library().lhip_memory_need_mode(_U64(n), METHOD_TREE if method == "tree" else METHOD_DIRECT,
                                             MODE_MIXED if mode == "mixed" else MODE_PLAIN, ctypes.byref(d),
                                             ctypes.byref(h))
        if rc != OK:
            raise ValueError("memory_need failed")
        return int(d.value), int(h.value)

    def device_bytes_held(self) -> int:
        """Device bytes this backend's force path holds (counted as free by device_memory)."""
        v = _U64()
        self._check(self._L.lhip_device_bytes_held(self._h, ctypes.byref(v)))
        return int(v.value)

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

    # ---- snapshots (src/snapshot.hpp; FP64 arrays in the mixed-precision mode) ----
    def write_snapshot(self, path: str, **header) -> None:
        """The bodies (download order) to an LTCSNAP3 file, with the FP64 state in the mixed mode."""
        c = _snap_header(header)
        self._check(self._L.lhip_snapshot_write(self._h, os.fsencode(path), ctypes.byref(c)))

    def restore_snapshot(self, path: str) -> dict:
        """A snapshot into this backend (as upload); the FP64 state when present and in the mixed mode."""
        c = SnapshotHeaderC()
        self._check(self._L.lhip_snapshot_restore(self._h, os.fsencode(path), ctypes.byref(c)))
        return _snap_dict(c)

    # ---- mixed-precision form (no reference counterpart) ----
    def set_precision(self, mode: str = "plain", group_max=None, small_leaf=None, quad_tau=None) -> None:
        """'plain': the reference's FP32 state; 'mixed': FP64 state with FP32 relative-coordinate
        pair arithmetic. Walk options (None: the mode's default, plain 128/4/0 = the reference walk,
        mixed 128/1/0.05): group_max (16..128), small_leaf (0..16), quad_tau (>= 0, the
        error-budgeted far field threshold)."""
        if mode not in ("plain", "mixed"):
            raise ValueError("mode must be 'plain' or 'mixed'")
        d = DEFAULT_WALK[mode]
        o = WalkOptionsC(int(d[0] if group_max is None else group_max), int(d[1] if small_leaf is None else small_leaf),
                         float(d[2] if quad_tau is None else quad_tau))
        self._check(self._L.lhip_set_precision(self._h, MODE_MIXED if mode == "mixed" else MODE_PLAIN,
                                               ctypes.byref(o)))

    def precision(self):
        """(mode, {group_max, small_leaf, quad_tau})."""
        m, o = ctypes.c_int32(), WalkOptionsC()
        self._check(self._L.lhip_get_precision(self._h, ctypes.byref(m), ctypes.byref(o)))
        return ("mixed" if m.value == MODE_MIXED else "plain"), dict(group_max=o.group_max, small_leaf=o.small_leaf,
                                                                     quad_tau=o.quad_tau)

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

    def direct_f64_targets(self, targets, G: float, eps: float):
        """FP64 direct-sum (a (k x 3), phi (k)) of the bodies at backend indices `targets`."""
        t = np.ascontiguousarray(targets, np.uint32)
        out = [np.zeros(len(t)) for _ in range(4)]
        self._check(self._L.lhip_direct_f64_targets(self._h, _U64(len(t)), _ptr(t), _D(G), _D(eps),
                                                    *[_ptr(o) for o in out]))
        return np.stack(out[:3], 1), out[3]

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
        return dict(lo=link[:, 0], hi=link[:, 1], level=link[:, 2] & 127, leaf=(link[:, 2] >> 7) & 1,
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
