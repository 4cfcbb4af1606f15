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

from . import HipBackend, LogicError, METHOD_DIRECT, OK, _D, _P, _U64, library

PRECISION_MIXED, PRECISION_F64 = 0, 1


class SimParamsC(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("eps", ctypes.c_double), ("dt", ctypes.c_double), ("theta", ctypes.c_double),
                ("steps", ctypes.c_int32), ("checkpoint_every", ctypes.c_int32), ("method", ctypes.c_int32),
                ("precision", ctypes.c_int32)]


@dataclass
class SimParams:
    G: float = 1.0
    eps: float = 0.05
    dt: float = 1e-3
    steps: int = 200
    checkpoint_every: int = 20
    precision: str = "mixed"  # "mixed" (FP32 pair arithmetic on FP64 separations) or "f64"

    def c(self) -> SimParamsC:
        if self.precision not in ("mixed", "f64"):
            raise ValueError("precision must be 'mixed' or 'f64'")
        return SimParamsC(self.G, self.eps, self.dt, 0.0, int(self.steps), int(self.checkpoint_every), METHOD_DIRECT,
                          PRECISION_F64 if self.precision == "f64" else PRECISION_MIXED)


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
        self.be._check(library().lhip_sim_forward(self.be._h, ctypes.byref(c), _U64(n), x0.ctypes.data_as(_P),
                                                  v0.ctypes.data_as(_P), m.ctypes.data_as(_P),
                                                  xT.ctypes.data_as(_P), vT.ctypes.data_as(_P)))
        self.n = n
        return xT, vT

    def backward(self, gxT, gvT):
        n = getattr(self, "n", None) or int(np.asarray(gxT).reshape(-1, 3).shape[0])
        gxT, gvT = _f64(gxT, (n, 3)), _f64(gvT, (n, 3))
        gx0, gv0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        self.be._check(library().lhip_sim_backward(self.be._h, gxT.ctypes.data_as(_P), gvT.ctypes.data_as(_P),
                                                   gx0.ctypes.data_as(_P), gv0.ctypes.data_as(_P),
                                                   gm.ctypes.data_as(_P)))
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
            ctx.sim = sim
            ctx.dev = x0.device
            return (torch.from_numpy(xT).to(x0.device), torch.from_numpy(vT).to(x0.device))

        @staticmethod
        def backward(ctx, gxT, gvT):
            n = ctx.sim.n
            gxT = np.zeros((n, 3)) if gxT is None else gxT.detach().cpu().numpy()
            gvT = np.zeros((n, 3)) if gvT is None else gvT.detach().cpu().numpy()
            gx0, gv0, gm = ctx.sim.backward(gxT, gvT)
            t = lambda a: torch.from_numpy(a).to(ctx.dev)  # noqa: E731
            return t(gx0), t(gv0), t(gm), None

    return Leapfrog


__all__ = ["SimParams", "Simulation", "leapfrog_function", "LogicError", "OK", "_D"]
