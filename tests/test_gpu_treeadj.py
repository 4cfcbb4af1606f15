"""GPU tests of the differentiable simulation through the tree (frozen-topology adjoint of the
mixed-precision walk): replay of a recorded forward, the adjoint against central differences of the
frozen-topology forward (parameters of the two-disk encounter and single coordinates), against the
direct-sum adjoint, and checkpoint invariance. No reference counterpart: the reference has no
differentiation."""
from __future__ import annotations

import numpy as np
import pytest

import lettuce_simulate_physics_amd as L
from lettuce_simulate_physics_amd.autograd import SimParams, Simulation
from helpers import restart_positions

pytestmark = pytest.mark.gpu

PRM = dict(method="tree", theta=0.6, eps=0.02, dt=5e-3, steps=40, checkpoint_every=8)


@pytest.fixture(autouse=True)
def plain_after(hip):
    # the tree simulation runs on the context's bodies in the mixed mode; restore the plain mode
    yield
    hip.set_precision("plain")


def encounter(nstars, seed=3):
    return L.Collision(nstars=nstars, seed=seed, pericentre=4.0, incl1=0.5, incl2=1.0)


def com2_loss(C, m, target):
    """Squared distance of disk 2's centre of mass from target: loss(xT) and dL/dxT."""
    n1 = C.nstars + 1
    w = np.zeros(len(m))
    w[n1:] = m[n1:] / m[n1:].sum()

    def f(xT):
        c = (w[:, None] * xT).sum(0)
        return float(((c - target) ** 2).sum()), 2 * w[:, None] * (c - target)[None, :]
    return f


def test_replay_reproduces_the_recorded_forward(hip):
    C = encounter(3000)
    x0, v0, m = C.generate()
    rec = Simulation(hip, SimParams(**{**PRM, "steps": 12}, tree_mode="record"))
    xa, va = rec.forward(x0, v0, m)
    rep = Simulation(hip, SimParams(**{**PRM, "steps": 12}, tree_mode="replay"))
    xb, vb = rep.forward(x0, v0, m)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    # and the plain tree forward is the same trajectory
    plain = Simulation(hip, SimParams(**{**PRM, "steps": 12}))
    xc, vc = plain.forward(x0, v0, m)
    assert np.array_equal(xa, xc) and np.array_equal(va, vc)


def test_replay_with_key_restarts(hip):
    # A recorded forward whose trees restart keys in over-full finest cells (a clump of 200 bodies
    # within 1e-6): the replay rebuilds each call's frames from the recorded positions and
    # reproduces the forward bit for bit; the adjoint runs on the same trees.
    x0 = restart_positions("clump").astype(np.float64)
    n = len(x0)
    v0 = np.random.default_rng(6).standard_normal((n, 3)) * 1e-3
    v0[:200] = 0.0  # the clump stays over-full
    m = np.full(n, 1.0 / n)
    prm = dict(method="tree", theta=0.6, eps=0.01, dt=1e-3, steps=6, checkpoint_every=3)
    rec = Simulation(hip, SimParams(**prm, tree_mode="record"))
    xa, va = rec.forward(x0, v0, m)
    assert int(hip.tree_nodes()["level"].max()) > 21
    gx, gv, _ = rec.backward(np.ones_like(xa), np.zeros_like(va))
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gv))
    rep = Simulation(hip, SimParams(**prm, tree_mode="replay"))
    xb, vb = rep.forward(x0, v0, m)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_tree_adjoint_parameter_gradient_matches_frozen_forward_fd(hip):
    # dL/d(pericentre, incl1, incl2) through the generator's VJP against central differences of the
    # frozen-topology forward (replay of the recorded topology with perturbed initial conditions).
    C = encounter(2000)
    x0, v0, m = C.generate()
    target = np.array([1.0, -2.0, 0.5])
    loss = com2_loss(C, m, target)
    rec = Simulation(hip, SimParams(**PRM, tree_mode="record"))
    xT, _ = rec.forward(x0, v0, m)
    L0, gxT = loss(xT)
    gx0, gv0, gm = rec.backward(gxT, np.zeros_like(gxT))
    assert gm is None
    g = C.vjp(gx0, gv0)
    rep = Simulation(hip, SimParams(**PRM, tree_mode="replay"))
    res = []
    for k, (name, h) in enumerate((("pericentre", 1e-3), ("incl1", 1e-3), ("incl2", 1e-3))):
        vals = []
        for sgn in (1, -1):
            Cp = L.Collision(**{**C.__dict__, name: getattr(C, name) + sgn * h})
            xp, vp, mp = Cp.generate()
            xq, _ = rep.forward(xp, vp, mp)
            vals.append(loss(xq)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        res.append((name, fd, g[k]))
    print("parameter gradients (FD, adjoint):", [(n, f"{a:.6g}", f"{b:.6g}") for n, a, b in res])
    for name, fd, ad in res:
        assert abs(fd - ad) <= 1e-4 * max(abs(fd), 1e-3 * max(abs(r[1]) for r in res)), name


@pytest.mark.parametrize("theta,walk", [
    (1e-6, dict(group_max=128, small_leaf=4, quad_tau=0.0)),   # every cell opened: body terms only
    (0.6, dict(group_max=128, small_leaf=4, quad_tau=0.0)),    # quadrupole cells, the reference walk
    (0.6, None),                                               # mixed defaults: monopole far field
    (0.6, dict(group_max=128, small_leaf=1, quad_tau=0.0)),    # accepted leaves as cells
], ids=["bodies", "quadrupoles", "monopoles", "small-leaf-1"])
def test_tree_adjoint_single_coordinates_match_frozen_forward_fd(hip, theta, walk):
    # The force Jacobian alone: a Plummer sphere, two steps, a loss on the final velocities, so each
    # gradient is the tidal / quadrupole / cell-moment terms of the VJP. Central differences of the
    # frozen-topology forward with Richardson extrapolation (h = 1e-3, 5e-4). The forward's pair
    # arithmetic is FP32, so the differences carry ~1e-4 of noise; the bound is 2e-3.
    b = L.plummer(4000, 3)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64)
    n = len(m)
    cv = np.random.default_rng(4).standard_normal((n, 3))
    prm = dict(method="tree", theta=theta, eps=0.05, dt=1e-2, steps=2, checkpoint_every=2, walk=walk)
    rec = Simulation(hip, SimParams(**prm, tree_mode="record"))
    xT, _ = rec.forward(x0, v0, m)
    gx, _, _ = rec.backward(np.zeros_like(xT), cv)
    rep = Simulation(hip, SimParams(**prm, tree_mode="replay"))

    def loss(x):
        _, w = rep.forward(x, v0, m)
        return float((cv * w).sum())

    rng = np.random.default_rng(5)
    scale = np.sqrt((gx ** 2).mean())
    worst = 0.0
    for _ in range(10):
        i, c = int(rng.integers(n)), int(rng.integers(3))
        fds = []
        for h in (1e-3, 5e-4):
            a, bb = x0.copy(), x0.copy()
            a[i, c] += h
            bb[i, c] -= h
            fds.append((loss(a) - loss(bb)) / (2 * h))
        fd = (4 * fds[1] - fds[0]) / 3
        worst = max(worst, abs(fd - gx[i, c]) / max(abs(fd), 0.1 * scale))
    print(f"single coordinates: worst relative {worst:.3g}")
    assert worst < 2e-3


BENCH_ENCOUNTER = dict(nstars=16384, seed=1, pericentre=4.0, incl1=np.pi / 6, incl2=np.pi / 3)
BENCH_PRM = dict(theta=0.6, eps=0.02, dt=1e-3, steps=100, checkpoint_every=20)
REF_WALK = dict(group_max=128, small_leaf=4, quad_tau=0.0)  # every quadrupole kept


def encounter_gradient(hip, method, walk=None, tree_mode="build"):
    """dL/d(pericentre, incl1, incl2) of the benchmark's encounter at 2 x 16384 stars."""
    C = L.Collision(**BENCH_ENCOUNTER)
    x0, v0, m = C.generate()
    loss = com2_loss(C, m, np.random.default_rng(2027).standard_normal(3))
    sim = Simulation(hip, SimParams(method=method, walk=walk, tree_mode=tree_mode, **BENCH_PRM))
    xT, _ = sim.forward(x0, v0, m)
    _, gxT = loss(xT)
    gx0, gv0, _ = sim.backward(gxT, np.zeros_like(gxT))
    return C, loss, C.vjp(gx0, gv0)


def test_tree_adjoint_agrees_with_the_direct_sum_adjoint(hip):
    # The two-disk encounter at 2 x 16384 stars with the benchmark's settings (inclinations 30 and 60
    # degrees, pericentre 4, theta 0.6, eps 0.02, dt 1e-3, 100 steps, checkpoints every 20) through the
    # tree and through the direct sum, per component. With every quadrupole kept (the reference walk)
    # the components agree to the tree's force accuracy: dL/d(incl1), which reaches disk 2 only through
    # disk 1's quadrupole and higher moments, to 3% (measured 3.1%), the others to 1e-3. With the mixed
    # form's error-budgeted far field (cells whose quadrupole term is below 5% of their monopole term
    # evaluated as point masses) dL/d(incl1) is about a third smaller: that far field drops the
    # quadrupoles that carry disk 1's orientation. The adjoint itself matches central differences of
    # its own frozen-topology forward in both walks (next test), so the gap is the force model's.
    _, _, direct = encounter_gradient(hip, "direct")
    _, _, full = encounter_gradient(hip, "tree", REF_WALK)
    _, _, budget = encounter_gradient(hip, "tree")
    rel = lambda a: np.abs(a - direct) / np.abs(direct)  # noqa: E731
    print(f"direct {direct}\ntree, every quadrupole {full} (relative {rel(full)})\n"
          f"tree, mixed far field {budget} (relative {rel(budget)})")
    assert rel(full)[0] < 1e-3 and rel(full)[2] < 1e-3 and rel(full)[1] < 0.05
    assert rel(budget)[0] < 1e-3 and rel(budget)[2] < 1e-3


@pytest.mark.parametrize("walk", [None, REF_WALK], ids=["mixed-far-field", "every-quadrupole"])
def test_tree_adjoint_matches_frozen_forward_fd_at_benchmark_size(hip, walk):
    # The same encounter: the tree adjoint's dL/d(pericentre, incl1, incl2) against central
    # differences of the frozen-topology forward (replays of the recorded topology from the
    # generator's perturbed initial conditions), Richardson-extrapolated from h = 1e-2 and 5e-3.
    # Per-component relative tolerance: 1e-5 (pericentre; measured 1e-7), 1e-4 (incl2; 2e-6) and 2e-2
    # (incl1, a gradient 1e5 times smaller than the pericentre's, where FP32 rounding in the forward
    # leaves ~0.5% of noise in the differences; measured 0.4% and 0.08%).
    C, loss, ad = encounter_gradient(hip, "tree", walk, tree_mode="record")
    rep = Simulation(hip, SimParams(method="tree", walk=walk, tree_mode="replay", **BENCH_PRM))
    fd = []
    for name in ("pericentre", "incl1", "incl2"):
        est = []
        for h in (1e-2, 5e-3):
            vals = []
            for sgn in (1, -1):
                xp, vp, mp = L.Collision(**{**BENCH_ENCOUNTER, name: getattr(C, name) + sgn * h}).generate()
                xq, _ = rep.forward(xp, vp, mp)
                vals.append(loss(xq)[0])
            est.append((vals[0] - vals[1]) / (2 * h))
        fd.append((4 * est[1] - est[0]) / 3)
    fd = np.array(fd)
    err = np.abs(ad - fd) / np.abs(fd)
    print(f"adjoint {ad}\nfrozen-forward FD {fd}\nrelative {err}")
    assert err[0] < 1e-5 and err[1] < 2e-2 and err[2] < 1e-4


def test_tree_checkpoint_interval_does_not_change_gradients(hip):
    C = encounter(1000, seed=7)
    x0, v0, m = C.generate()
    out = []
    for K in (1, 5, 40):
        sim = Simulation(hip, SimParams(**{**PRM, "checkpoint_every": K}))
        xT, vT = sim.forward(x0, v0, m)
        out.append(sim.backward(np.ones_like(xT), vT)[:2])
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            # source-side sums arrive by atomic adds: equal up to summation order
            assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))


def test_backward_after_a_replay_is_the_recorded_forwards_adjoint(hip):
    # A replay (other initial conditions, other step count per checkpoint) leaves the recorded
    # forward's checkpoints and parameters in place: the recording's backward after it returns the
    # same gradients as before it (up to the order of the atomic source-side sums); the replay
    # itself has no adjoint of its own, and one with another theta or walk options is refused.
    C = encounter(1500, seed=11)
    x0, v0, m = C.generate()
    rec = Simulation(hip, SimParams(**{**PRM, "steps": 8}, tree_mode="record"))
    xT, vT = rec.forward(x0, v0, m)
    g1 = rec.backward(xT, vT)[:2]
    rep = Simulation(hip, SimParams(**{**PRM, "steps": 8, "checkpoint_every": 3, "dt": 4e-3}, tree_mode="replay"))
    rep.forward(x0 * 1.001, v0, m)
    with pytest.raises(L.LogicError):
        rep.backward(xT, vT)
    g2 = rec.backward(xT, vT)[:2]
    for a, b in zip(g2, g1):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))
    for bad in (dict(theta=0.5), dict(walk=dict(group_max=64, small_leaf=1, quad_tau=0.05))):
        with pytest.raises(L.LogicError):
            Simulation(hip, SimParams(**{**PRM, "steps": 8, **bad}, tree_mode="replay")).forward(x0, v0, m)
    g3 = rec.backward(xT, vT)[:2]
    for a, b in zip(g3, g1):
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))


def test_tree_backward_without_mass_gradient_and_replay_errors(hip):
    C = encounter(500, seed=9)
    x0, v0, m = C.generate()
    sim = Simulation(hip, SimParams(**{**PRM, "steps": 4}))
    xT, vT = sim.forward(x0, v0, m)
    lib = L.library()
    import ctypes
    n = len(m)
    z = np.zeros((n, 3))
    gm = np.zeros(n)
    rc = lib.lhip_sim_backward(hip._h, z.ctypes.data_as(ctypes.c_void_p), z.ctypes.data_as(ctypes.c_void_p),
                               None, None, gm.ctypes.data_as(ctypes.c_void_p))
    assert rc == L.INVALID
    other = encounter(600, seed=9)
    x1, v1, m1 = other.generate()
    with pytest.raises(L.LogicError):
        Simulation(hip, SimParams(**{**PRM, "steps": 4}, tree_mode="replay")).forward(x1, v1, m1)
