"""Gradient harness: the differentiable FP64-state KDK simulation of liblettuce_hip.so as a
``torch.autograd.Function``. PyTorch only carries the gradients; forward and adjoint run in the HIP
library (``lhip_sim_forward`` / ``lhip_sim_backward``).

No reference counterpart: the reference has no differentiation. The forward integrates the
reference's KDK leapfrog (src/simulation.cpp:406-438) with Plummer-softened direct forces.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import (DEFAULT_WALK, HipBackend, LogicError, METHOD_DIRECT, METHOD_TREE, OK, WalkOptionsC, _D, _P, _U64,
               library)

PRECISION_MIXED, PRECISION_F64 = 0, 1
TREE_MODES = {"build": 0, "record": 1, "replay": 2}


class SimParamsC(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("eps", ctypes.c_double), ("dt", ctypes.c_double), ("theta", ctypes.c_double),
                ("steps", ctypes.c_int32), ("checkpoint_every", ctypes.c_int32), ("method", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("tree_mode", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("walk", WalkOptionsC)]


@dataclass
class SimParams:
    G: float = 1.0
    eps: float = 0.05
    dt: float = 1e-3
    steps: int = 200
    checkpoint_every: int = 20
    precision: str = "mixed"  # direct: "mixed" (FP32 pair arithmetic on FP64 separations) or "f64"
    method: str = "direct"    # or "tree": the mixed-precision walk with frozen topology in the adjoint
    theta: float = 0.6
    tree_mode: str = "build"  # tree: "build", "record" (keep each step's topology) or "replay"
    walk: dict | None = None  # tree walk options (group_max, small_leaf, quad_tau); None: mixed defaults

    def c(self) -> SimParamsC:
        if self.precision not in ("mixed", "f64"):
            raise ValueError("precision must be 'mixed' or 'f64'")
        if self.method not in ("direct", "tree"):
            raise ValueError("method must be 'direct' or 'tree'")
        w = WalkOptionsC(0, 0, 0.0)
        if self.walk is not None:
            d = DEFAULT_WALK["mixed"]
            w = WalkOptionsC(int(self.walk.get("group_max", d[0])), int(self.walk.get("small_leaf", d[1])),
                             float(self.walk.get("quad_tau", d[2])))
        return SimParamsC(self.G, self.eps, self.dt, self.theta, int(self.steps), int(self.checkpoint_every),
                          METHOD_TREE if self.method == "tree" else METHOD_DIRECT,
                          PRECISION_F64 if self.precision == "f64" else PRECISION_MIXED, TREE_MODES[self.tree_mode], 0,
                          w)


def _f64(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))


class Simulation:
    """Forward with checkpoints and adjoint on one HIP backend context."""

    def __init__(self, backend: HipBackend, params: SimParams):
        self.be = backend
        self.p = params

    def forward(self, x0, v0, m):
        n = len(m)
        x0, v0, m = _f64(x0, (n, 3)), _f64(v0, (n, 3)), _f64(m, (n,))
        xT, vT = np.zeros((n, 3)), np.zeros((n, 3))
        c = self.p.c()
        self.gen = None
        self.be._check(library().lhip_sim_forward(self.be._h, ctypes.byref(c), _U64(n), x0.ctypes.data_as(_P),
                                                  v0.ctypes.data_as(_P), m.ctypes.data_as(_P),
                                                  xT.ctypes.data_as(_P), vT.ctypes.data_as(_P)))
        self.n = n
        if self.p.method != "tree" or self.p.tree_mode != "replay":
            gen = ctypes.c_uint64()
            self.be._check(library().lhip_sim_generation(self.be._h, ctypes.byref(gen)))
            self.gen = gen.value  # the forward whose checkpoints backward() needs
        return xT, vT

    def backward(self, gxT, gvT):
        """(dL/dx0, dL/dv0, dL/dm) of this object's last forward; dL/dm is None for the tree method
        (masses are not differentiated through the tree). Raises LogicError when another forward has
        run on the same backend since (its checkpoints are gone), and for a replay, which has no
        adjoint of its own (the recording Simulation's backward is the frozen-topology adjoint)."""
        return self.backward_of(getattr(self, "gen", None), gxT, gvT)

    def backward_of(self, gen, gxT, gvT):
        """backward() for the forward of generation `gen` (what forward() left in self.gen)."""
        if gen is None:
            raise LogicError("sim backward: no forward of this Simulation to differentiate "
                             "(none ran, or it was a replay)")
        n = int(np.asarray(gxT).reshape(-1, 3).shape[0])
        gxT, gvT = _f64(gxT, (n, 3)), _f64(gvT, (n, 3))
        gx0, gv0 = np.zeros((n, 3)), np.zeros((n, 3))
        gm = None if self.p.method == "tree" else np.zeros(n)
        self.be._check(library().lhip_sim_backward_gen(self.be._h, _U64(gen), gxT.ctypes.data_as(_P),
                                                       gvT.ctypes.data_as(_P), gx0.ctypes.data_as(_P),
                                                       gv0.ctypes.data_as(_P),
                                                       None if gm is None else gm.ctypes.data_as(_P)))
        return gx0, gv0, gm

    def times(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        self.be._check(library().lhip_sim_times(self.be._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value


def leapfrog_function():
    """The torch.autograd.Function class (imported lazily so the package does not need torch)."""
    import torch

    class Leapfrog(torch.autograd.Function):
        """(x0, v0, m) -> (xT, vT) through `sim`; backward runs the HIP adjoint sweep."""

        @staticmethod
        def forward(ctx, x0, v0, m, sim: Simulation):
            xT, vT = sim.forward(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), m.detach().cpu().numpy())
            # This graph's forward: a later forward on the same backend makes its backward raise
            # rather than return another trajectory's gradients.
            ctx.sim, ctx.gen, ctx.n = sim, sim.gen, sim.n
            ctx.dev = x0.device
            return (torch.from_numpy(xT).to(x0.device), torch.from_numpy(vT).to(x0.device))

        @staticmethod
        def backward(ctx, gxT, gvT):
            n = ctx.n
            gxT = np.zeros((n, 3)) if gxT is None else gxT.detach().cpu().numpy()
            gvT = np.zeros((n, 3)) if gvT is None else gvT.detach().cpu().numpy()
            gx0, gv0, gm = ctx.sim.backward_of(ctx.gen, gxT, gvT)
            t = lambda a: None if a is None else torch.from_numpy(a).to(ctx.dev)  # noqa: E731
            return t(gx0), t(gv0), t(gm), None

    return Leapfrog


__all__ = ["SimParams", "Simulation", "leapfrog_function", "LogicError", "OK", "_D"]
