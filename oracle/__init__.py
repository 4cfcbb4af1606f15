"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline).

ctypes wrappers for two CPU libraries built by ``oracle/Makefile``:

* ``liboracle.so`` — this repository's restatement of the reference CPU force path
  (``oracle/lettuce_oracle.cpp``; each function there cites the reference file:line it follows);
* ``_ref/libref.so`` — the reference's own CPU backend and scenario generator compiled from its
  sources (present only where it was built; optional everywhere).

The product path (``lettuce_simulate_physics_amd``) never imports this package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_D = ctypes.c_double
_U64 = ctypes.c_uint64


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and _ref/libref.so when ref and the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_create.restype = _P
        L.orc_size.restype = _U64
        L.orc_exact_pe.restype = _D
        L.orc_node_count.restype = _U64
        L.orc_group_count.restype = _U64
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref():
    """The reference CPU backend library, or None where it was not built."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(os.path.join(HERE, "_ref", "libref.so"))
        R.ref_scenario.restype = ctypes.c_int64
        R.ref_cpu_create.restype = _P
        R.ref_exact_pe.restype = _D
        R.ref_last_error.restype = ctypes.c_char_p
        _ref = R
    return _ref


@dataclass
class Bodies:
    """Host SoA bodies, as the reference's Bodies (src/bodies.hpp:8-13)."""
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
        f = [np.zeros(n, np.float32) for _ in range(7)]
        return Bodies(*f, np.zeros(n, np.uint32))

    def arrays(self):
        return [self.x, self.y, self.z, self.vx, self.vy, self.vz, self.m, self.id]

    def __len__(self):
        return len(self.x)

    def copy(self) -> "Bodies":
        return Bodies(*[a.copy() for a in self.arrays()])

    def subset(self, idx) -> "Bodies":
        return Bodies(*[np.ascontiguousarray(a[idx]) for a in self.arrays()])


def plummer(n: int, seed: int = 1, a: float = 1.0, cut: float = 10.0, M: float = 1.0, G: float = 1.0) -> Bodies:
    b = Bodies.empty(n)
    lib().orc_plummer(_U64(n), _U64(seed), _D(G), _D(M), _D(a), _D(cut), *[_ptr(v) for v in b.arrays()])
    return b


def disk_sech2(nstars: int, seed: int = 1, core: float = 10.0, mass: float = 1.0, rd: float = 1.0,
               rmax: float = 10.0, z0: float = 0.1, sigma: float = 0.1, incl: float = 0.0, spin: float = 1.0,
               pos=(0.0, 0.0, 0.0), vel=(0.0, 0.0, 0.0), com_frame: bool = True, G: float = 1.0) -> Bodies:
    b = Bodies.empty(nstars + 1)
    lib().orc_disk_sech2(_U64(nstars), _U64(seed), _D(G), _D(core), _D(mass), _D(rd), _D(rmax), _D(z0), _D(sigma),
                         _D(incl), _D(spin), *[_D(v) for v in pos], *[_D(v) for v in vel], int(com_frame),
                         *[_ptr(v) for v in b.arrays()])
    return b


def ref_scenario(name: str, n: int, seed: int = 1, params: str = "", G: float = 1.0) -> Bodies:
    R = ref()
    if R is None:
        raise RuntimeError("reference library not built (oracle/_ref/libref.so)")
    cnt = R.ref_scenario(name.encode(), _U64(n), _U64(seed), _D(G), params.encode(), *([None] * 8))
    if cnt < 0:
        raise RuntimeError(R.ref_last_error().decode())
    b = Bodies.empty(int(cnt))
    got = R.ref_scenario(name.encode(), _U64(n), _U64(seed), _D(G), params.encode(), *[_ptr(v) for v in b.arrays()])
    if got != cnt:
        raise RuntimeError(R.ref_last_error().decode())
    return b


def direct_f64(b: Bodies, G: float, eps: float, targets=None):
    """FP64 brute-force accelerations and potentials for target indices (default: all)."""
    n = len(b)
    idx = np.arange(n, dtype=np.int64) if targets is None else np.ascontiguousarray(targets, dtype=np.int64)
    pos = [np.ascontiguousarray(v, dtype=np.float64) for v in (b.x, b.y, b.z, b.m)]
    out = [np.zeros(len(idx)) for _ in range(4)]
    lib().orc_direct_f64(_U64(n), *[_ptr(p) for p in pos], _D(G), _D(eps), _U64(len(idx)), _ptr(idx),
                         *[_ptr(o) for o in out])
    return np.stack(out[:3], axis=1), out[3]


def sim_forward(x0, v0, m, G, eps, dt, steps):
    """FP64 KDK with FP64 direct forces (the oracle's differentiable simulation); (xT, vT)."""
    n = len(m)
    x0 = np.ascontiguousarray(x0, np.float64).reshape(n, 3)
    v0 = np.ascontiguousarray(v0, np.float64).reshape(n, 3)
    m = np.ascontiguousarray(m, np.float64)
    xT, vT = np.zeros((n, 3)), np.zeros((n, 3))
    lib().orc_sim_forward(_U64(n), _ptr(x0), _ptr(v0), _ptr(m), _D(G), _D(eps), _D(dt), int(steps), _ptr(xT), _ptr(vT))
    return xT, vT


def sim_backward(gxT, gvT):
    """Adjoint of the last sim_forward: (dL/dx0, dL/dv0, dL/dm)."""
    gxT = np.ascontiguousarray(gxT, np.float64)
    gvT = np.ascontiguousarray(gvT, np.float64)
    n = gxT.shape[0]
    gx0, gv0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_sim_backward(_ptr(gxT), _ptr(gvT), _ptr(gx0), _ptr(gv0), _ptr(gm))
    return gx0, gv0, gm


# LetNode (src/backend.hpp:63-76): the 60-byte wire record of a locally essential tree.
LET_NODE = np.dtype([(k, "<f4") for k in ("cx", "cy", "cz", "mass", "qxx", "qxy", "qxz", "qyy", "qyz", "qzz",
                                          "rcrit2")] +
                    [(k, "<i4") for k in ("first_child", "child_count", "body_first", "body_count")])


@dataclass
class Let:
    """LetTree (src/backend.hpp:84-87): LET_NODE records and the exported bodies."""
    nodes: np.ndarray
    bx: np.ndarray
    by: np.ndarray
    bz: np.ndarray
    bm: np.ndarray

    @staticmethod
    def sized(nn: int, nb: int) -> "Let":
        return Let(np.zeros(nn, LET_NODE), *[np.zeros(nb, np.float32) for _ in range(4)])


class _FlatLet(ctypes.Structure):
    _fields_ = [("nodes", _P), ("n_nodes", ctypes.c_uint64), ("bx", _P), ("by", _P), ("bz", _P), ("bm", _P),
                ("n_bodies", ctypes.c_uint64)]


def _flat(remote):
    """(ctypes array of flat LETs, arrays to keep alive) for objects with nodes, bx, by, bz, bm."""
    keep, arr = [], (_FlatLet * max(len(remote), 1))()
    for k, t in enumerate(remote):
        nodes = np.ascontiguousarray(t.nodes, LET_NODE)
        cols = [np.ascontiguousarray(v, np.float32) for v in (t.bx, t.by, t.bz, t.bm)]
        keep += [nodes, *cols]
        arr[k] = _FlatLet(_ptr(nodes), len(nodes), *[_ptr(v) for v in cols], len(cols[0]))
    return arr, keep


def ref_write_snapshot(path: str, header: dict, b: Bodies) -> None:
    """The reference's write_snapshot (src/snapshot.hpp:23), from oracle/_ref."""
    R = ref()
    hd = np.zeros(8, np.float64)
    u = hd.view(np.uint64)
    u[1], u[6], u[7] = header.get("step", 0), header.get("seed", 0), header.get("pieces", 0)
    hd[2], hd[3], hd[4], hd[5] = header.get("time", 0.0), header.get("G", 1.0), header.get("eps", 0.0), header.get("dt", 0.0)
    if R.ref_write_snapshot(os.fsencode(path), _ptr(hd), str(header.get("scenario", "")).encode(), _U64(len(b)),
                            *[_ptr(v) for v in b.arrays()]):
        raise RuntimeError(R.ref_last_error().decode())


def ref_read_snapshot(path: str):
    """The reference's read_snapshot (src/snapshot.hpp:24): (header, Bodies)."""
    R = ref()
    hd = np.zeros(8, np.float64)
    sc = ctypes.create_string_buffer(32)
    if R.ref_read_snapshot(os.fsencode(path), _ptr(hd), sc):
        raise RuntimeError(R.ref_last_error().decode())
    u = hd.view(np.uint64)
    h = dict(n=int(u[0]), step=int(u[1]), time=float(hd[2]), G=float(hd[3]), eps=float(hd[4]), dt=float(hd[5]),
             seed=int(u[6]), pieces=int(u[7]), scenario=sc.value.decode())
    b = Bodies.empty(h["n"])
    R.ref_snapshot_copy(*[_ptr(v) for v in b.arrays()])
    return h, b


class _Common:
    """Accelerations are read as the reference tests do: upload with v = 0, compute, kick(1), read v
    (tests/test_cpu_backend.cpp:45-60)."""

    def accel_by_id(self, b: Bodies, G, eps, theta, method, potential=False):
        z = b.copy()
        z.vx[:] = 0
        z.vy[:] = 0
        z.vz[:] = 0
        self.upload(z)
        self.compute_forces(G, eps, theta, method, potential)
        self.kick(1.0)
        o = self.download()
        a = np.zeros((len(o), 3))
        a[o.id, 0] = o.vx
        a[o.id, 1] = o.vy
        a[o.id, 2] = o.vz
        return a


class Oracle(_Common):
    """The restated CPU backend (oracle/lettuce_oracle.cpp)."""

    def __init__(self, threads: int = 0):
        self.L = lib()
        self.h = _P(self.L.orc_create(threads))

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def size(self) -> int:
        return int(self.L.orc_size(self.h))

    def upload(self, b: Bodies):
        self.L.orc_upload(self.h, _U64(len(b)), *[_ptr(v) for v in b.arrays()])

    def download(self) -> Bodies:
        b = Bodies.empty(self.size())
        self.L.orc_download(self.h, *[_ptr(v) for v in b.arrays()])
        return b

    def compute_forces(self, G=1.0, eps=0.01, theta=0.6, method=1, potential=False):
        if self.L.orc_compute_forces(self.h, _D(G), _D(eps), _D(theta), int(method), int(potential)):
            raise RuntimeError(self.L.orc_last_error().decode())

    def kick(self, dt):
        self.L.orc_kick(self.h, _D(dt))

    def drift(self, dt):
        self.L.orc_drift(self.h, _D(dt))

    def diagnostics(self) -> dict:
        d = (ctypes.c_double * 13)()
        self.L.orc_diagnostics(self.h, d)
        keys = ["mass", "ke", "pe", "has_pe", "px", "py", "pz", "lx", "ly", "lz", "cx", "cy", "cz"]
        return dict(zip(keys, list(d)))

    def exact_pe(self, G, eps) -> float:
        return float(self.L.orc_exact_pe(self.h, _D(G), _D(eps)))

    def counts(self) -> dict:
        c = (ctypes.c_uint64 * 4)()
        self.L.orc_counts(self.h, c)
        return dict(interactions=c[0], listed=c[1], p2p=c[2], p2c=c[3])

    def accel(self):
        n = self.size()
        out = [np.zeros(n, np.float32) for _ in range(4)]
        self.L.orc_accel(self.h, *[_ptr(o) for o in out])
        return out

    def body_costs(self):
        c = np.zeros(self.size(), np.float32)
        return c if self.L.orc_body_costs(self.h, _ptr(c)) else None

    def keys(self):
        n = self.size()
        k = np.zeros(n, np.uint64)
        p = np.zeros(n, np.uint32)
        self.L.orc_keys(self.h, _ptr(k), _ptr(p))
        return k, p

    def cube(self):
        c = (ctypes.c_double * 4)()
        self.L.orc_cube(self.h, c)
        return list(c)

    def nodes(self) -> dict:
        k = int(self.L.orc_node_count(self.h))
        ints = np.zeros((k, 8), np.int32)
        pre = np.zeros(k, np.uint64)
        flt = np.zeros((k, 16), np.float32)
        self.L.orc_nodes(self.h, _ptr(ints), _ptr(pre), _ptr(flt))
        return dict(level=ints[:, 0], depth=ints[:, 1], frame=ints[:, 2], lo=ints[:, 3], hi=ints[:, 4],
                    nchild=ints[:, 5], first=ints[:, 6], prefix=pre, com=flt[:, 0:3], m=flt[:, 3],
                    q=flt[:, 4:10], rcrit2=flt[:, 10], geo=flt[:, 11:15])

    def groups(self):
        k = int(self.L.orc_group_count(self.h))
        g = np.zeros((k, 2), np.uint32)
        self.L.orc_groups(self.h, _ptr(g))
        return g

    # ---- distributed support (src/backend.hpp:114-136) ----
    def local_bounds(self):
        lo, hi = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
        self.L.orc_local_bounds(self.h, lo, hi)
        return list(lo), list(hi)

    def build_tree(self, theta=0.6, method=1):
        self.L.orc_build_tree(self.h, _D(theta), int(method))

    def export_let(self, box, method=1) -> Let:
        lo = (ctypes.c_float * 3)(*[float(v) for v in box[0]])
        hi = (ctypes.c_float * 3)(*[float(v) for v in box[1]])
        nn, nb = _U64(), _U64()
        if self.L.orc_export_let(self.h, lo, hi, int(method), ctypes.byref(nn), ctypes.byref(nb)):
            raise RuntimeError(self.L.orc_last_error().decode())
        t = Let.sized(nn.value, nb.value)
        self.L.orc_export_let_copy(self.h, _ptr(t.nodes), _ptr(t.bx), _ptr(t.by), _ptr(t.bz), _ptr(t.bm))
        return t

    def compute_forces_with(self, G=1.0, eps=0.01, method=1, potential=False, remote=()):
        arr, _keep = _flat(remote)
        if self.L.orc_forces_with(self.h, _D(G), _D(eps), int(method), int(potential), arr, len(remote)):
            raise RuntimeError(self.L.orc_last_error().decode())


class RefCpu(_Common):
    """The reference's own CPU backend (oracle/_ref/libref.so), where it was built."""

    def __init__(self, threads: int = 0):
        self.R = ref()
        if self.R is None:
            raise RuntimeError("reference library not built (oracle/_ref/libref.so)")
        self.h = _P(self.R.ref_cpu_create(threads))
        self.n = 0

    def __del__(self):
        try:
            self.R.ref_cpu_destroy(self.h)
        except Exception:
            pass

    def upload(self, b: Bodies):
        self.n = len(b)
        self.R.ref_upload(self.h, _U64(len(b)), *[_ptr(v) for v in b.arrays()])

    def download(self) -> Bodies:
        b = Bodies.empty(self.n)
        self.R.ref_download(self.h, *[_ptr(v) for v in b.arrays()])
        return b

    def compute_forces(self, G=1.0, eps=0.01, theta=0.6, method=1, potential=False):
        if self.R.ref_compute_forces(self.h, _D(G), _D(eps), _D(theta), int(method), int(potential)):
            raise RuntimeError(self.R.ref_last_error().decode())

    def kick(self, dt):
        self.R.ref_kick(self.h, _D(dt))

    def drift(self, dt):
        self.R.ref_drift(self.h, _D(dt))

    def diagnostics(self) -> dict:
        d = (ctypes.c_double * 13)()
        self.R.ref_diagnostics(self.h, d)
        keys = ["mass", "ke", "pe", "has_pe", "px", "py", "pz", "lx", "ly", "lz", "cx", "cy", "cz"]
        return dict(zip(keys, list(d)))

    def exact_pe(self, G, eps) -> float:
        return float(self.R.ref_exact_pe(self.h, _D(G), _D(eps)))

    def times(self) -> dict:
        t = (ctypes.c_double * 5)()
        i = ctypes.c_uint64()
        self.R.ref_times(self.h, t, ctypes.byref(i))
        return dict(force_ms=t[0], sort_ms=t[1], build_ms=t[2], walk_ms=t[3], step_ms=t[4], interactions=i.value)

    def body_costs(self):
        c = np.zeros(self.n, np.float32)
        return c if self.R.ref_body_costs(self.h, _ptr(c)) else None

    # ---- distributed support (src/backend.hpp:114-136) ----
    def local_bounds(self):
        lo, hi = (ctypes.c_float * 3)(), (ctypes.c_float * 3)()
        self.R.ref_local_bounds(self.h, lo, hi)
        return list(lo), list(hi)

    def build_tree(self, G=1.0, eps=0.01, theta=0.6, method=1):
        if self.R.ref_build_tree(self.h, _D(G), _D(eps), _D(theta), int(method)):
            raise RuntimeError(self.R.ref_last_error().decode())

    def export_let(self, box, G=1.0, eps=0.01, theta=0.6, method=1) -> Let:
        lo = (ctypes.c_float * 3)(*[float(v) for v in box[0]])
        hi = (ctypes.c_float * 3)(*[float(v) for v in box[1]])
        nn, nb = _U64(), _U64()
        if self.R.ref_export_let(self.h, lo, hi, _D(G), _D(eps), _D(theta), int(method), ctypes.byref(nn),
                                 ctypes.byref(nb)):
            raise RuntimeError(self.R.ref_last_error().decode())
        t = Let.sized(nn.value, nb.value)
        self.R.ref_export_let_copy(_ptr(t.nodes), _ptr(t.bx), _ptr(t.by), _ptr(t.bz), _ptr(t.bm))
        return t

    def compute_forces_with(self, G=1.0, eps=0.01, theta=0.6, method=1, potential=False, remote=()):
        arr, _keep = _flat(remote)
        if self.R.ref_forces_with(self.h, _D(G), _D(eps), _D(theta), int(method), int(potential), arr, len(remote)):
            raise RuntimeError(self.R.ref_last_error().decode())

    def export_tree(self, G=1.0, eps=0.01, theta=0.6) -> dict:
        """The whole octree, breadth first (build_tree + export_let over a box holding every body)."""
        nn, nb = _U64(), _U64()
        if self.R.ref_export_tree(self.h, _D(G), _D(eps), _D(theta), ctypes.byref(nn), ctypes.byref(nb),
                                  None, None, None, None, None):
            raise RuntimeError(self.R.ref_last_error().decode())
        raw = np.zeros((nn.value, 15), np.float32)
        bod = [np.zeros(nb.value, np.float32) for _ in range(4)]
        self.R.ref_export_tree(self.h, _D(G), _D(eps), _D(theta), ctypes.byref(nn), ctypes.byref(nb), _ptr(raw),
                               *[_ptr(v) for v in bod])
        ints = raw[:, 11:15].view(np.int32)
        return dict(com=raw[:, 0:3], m=raw[:, 3], q=raw[:, 4:10], rcrit2=raw[:, 10], first_child=ints[:, 0],
                    child_count=ints[:, 1], body_first=ints[:, 2], body_count=ints[:, 3], bx=bod[0], by=bod[1],
                    bz=bod[2], bm=bod[3])
