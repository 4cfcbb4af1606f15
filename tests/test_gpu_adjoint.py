"""GPU tests of the differentiable simulation (lhip_sim_*): forward against the oracle's FP64
integrator, adjoint against the oracle's FP64 CPU adjoint and against central finite differences of
the all-FP64 GPU forward, checkpoint invariance, and the torch.autograd.Function harness."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O
from lettuce_simulate_physics_amd.autograd import SimParams, Simulation, leapfrog_function

pytestmark = pytest.mark.gpu


def plummer_state(n, seed, mass_noise=0.01):
    b = O.plummer(n, seed)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64) * (1 + mass_noise * np.random.default_rng(seed).standard_normal(n))
    return x0, v0, m


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_forward_matches_oracle_fp64(hip):
    x0, v0, m = plummer_state(2048, 3)
    for prec, tol in (("mixed", 1e-6), ("f64", 1e-12)):
        sim = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=20, checkpoint_every=5, precision=prec))
        xT, vT = sim.forward(x0, v0, m)
        xo, vo = O.sim_forward(x0, v0, m, 1.0, 0.05, 1e-3, 20)
        assert rel(xT - x0, xo - x0) < tol, prec
        assert rel(vT, vo) < tol * 10, prec


def test_adjoint_matches_oracle_fp64_adjoint(hip):
    x0, v0, m = plummer_state(2048, 4)
    tgt = x0 + 0.01 * np.random.default_rng(9).standard_normal(x0.shape)
    sim = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=20, checkpoint_every=7))
    xT, vT = sim.forward(x0, v0, m)
    gx, gv, gm = sim.backward(2 * (xT - tgt), np.zeros_like(vT))
    xo, vo = O.sim_forward(x0, v0, m, 1.0, 0.05, 1e-3, 20)
    ox, ov, om = O.sim_backward(2 * (xo - tgt), np.zeros_like(vo))
    print(f"adjoint vs FP64 CPU adjoint: x {rel(gx, ox):.3g} v {rel(gv, ov):.3g} m {rel(gm, om):.3g}")
    assert rel(gx, ox) < 1e-5 and rel(gv, ov) < 1e-5 and rel(gm, om) < 1e-4


def test_adjoint_matches_central_finite_differences(hip):
    # Central differences of the all-FP64 GPU forward over one crossing time (200 steps of 5e-3), where
    # the force Jacobian carries a large share of every gradient. The mixed-precision adjoint must agree
    # to 1e-4 relative per coordinate, the FP64 adjoint to 1e-6 (the differences themselves are good
    # to ~1e-7 at h = 1e-5); the floor for near-zero coordinates is a tenth of that kind's rms gradient.
    x0, v0, m = plummer_state(512, 5)
    tgt = x0 + 0.1 * np.random.default_rng(11).standard_normal(x0.shape)
    prm = dict(eps=0.05, dt=5e-3, steps=200, checkpoint_every=16)
    f64 = Simulation(hip, SimParams(**prm, precision="f64"))

    def loss(x, v, mm):
        a, w = f64.forward(x, v, mm)
        return float(((a - tgt) ** 2).sum() + (w ** 2).sum() / 3)

    grads = {}
    for prec in ("mixed", "f64"):
        sim = Simulation(hip, SimParams(**prm, precision=prec))
        xT, vT = sim.forward(x0, v0, m)
        grads[prec] = sim.backward(2 * (xT - tgt), 2 * vT / 3)
    gx = grads["f64"][0]
    assert np.sqrt(((gx - 2 * (xT - tgt)) ** 2).mean()) > 0.3 * np.sqrt((gx ** 2).mean())

    h = 1e-5
    rng = np.random.default_rng(3)
    worst = {"mixed": 0.0, "f64": 0.0}
    for _ in range(6):
        i, c = int(rng.integers(len(m))), int(rng.integers(3))
        for kind in range(3):
            a, b = [(x0, v0, m) for _ in range(2)]
            if kind < 2:
                a, b = list(a), list(b)
                a[kind], b[kind] = a[kind].copy(), b[kind].copy()
                a[kind][i, c] += h
                b[kind][i, c] -= h
                fd = (loss(*a) - loss(*b)) / (2 * h)
            else:
                mp, mm = m.copy(), m.copy()
                mp[i] += h * m[i]
                mm[i] -= h * m[i]
                fd = (loss(x0, v0, mp) - loss(x0, v0, mm)) / (2 * h * m[i])
            for prec, g in grads.items():
                gk = g[kind]
                val = gk[i, c] if kind < 2 else gk[i]
                floor = 0.1 * np.sqrt((gk ** 2).mean())
                worst[prec] = max(worst[prec], abs(fd - val) / max(floor, abs(fd)))
    print(f"adjoint vs central FD (18 coordinates): worst relative mixed {worst['mixed']:.3g}, "
          f"f64 {worst['f64']:.3g}")
    assert worst["mixed"] <= 1e-4
    assert worst["f64"] <= 1e-6


def test_adjoint_fd_on_64_coordinates_of_the_65536_body_run_shape(hip):
    # The adjoint benchmark's problem at N = 4096: Plummer (seed 1), eps 0.05, 200 KDK steps of 1e-3,
    # checkpoints every 20; target = the forward run from positions perturbed by N(0, 0.01^2) and masses
    # scaled by (1 + 0.01 N(0, 1)); loss = sum |x_T - x_target|^2. 64 random coordinates (positions,
    # velocities, masses) against central differences of the FP64 GPU forward, to 1e-4 relative (floor
    # for near-zero coordinates: a tenth of that kind's rms gradient).
    n, prm = 4096, dict(eps=0.05, dt=1e-3, steps=200, checkpoint_every=20)
    b = O.plummer(n, 1)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64)
    rng = np.random.default_rng(2026)
    f64 = Simulation(hip, SimParams(**prm, precision="f64"))
    xt, _ = f64.forward(x0 + 0.01 * rng.standard_normal(x0.shape), v0, m * (1 + 0.01 * rng.standard_normal(n)))

    def loss(x, v, mm):
        a, _ = f64.forward(x, v, mm)
        return float(((a - xt) ** 2).sum())

    sim = Simulation(hip, SimParams(**prm))
    xT, vT = sim.forward(x0, v0, m)
    g = sim.backward(2 * (xT - xt), np.zeros_like(vT))
    pick = np.random.default_rng(64)
    worst = [0.0, 0.0, 0.0]
    for k in range(64):
        kind, i, c = k % 3, int(pick.integers(n)), int(pick.integers(3))
        if kind < 2:
            h = 1e-5
            a, bb = [x0, v0, m], [x0, v0, m]
            a[kind], bb[kind] = a[kind].copy(), bb[kind].copy()
            a[kind][i, c] += h
            bb[kind][i, c] -= h
            fd, val = (loss(*a) - loss(*bb)) / (2 * h), g[kind][i, c]
        else:
            dm = 1e-4 * m[i]
            mp, mm = m.copy(), m.copy()
            mp[i] += dm
            mm[i] -= dm
            fd, val = (loss(x0, v0, mp) - loss(x0, v0, mm)) / (2 * dm), g[2][i]
        floor = 0.1 * np.sqrt((g[kind] ** 2).mean())
        worst[kind] = max(worst[kind], abs(fd - val) / max(floor, abs(fd)))
    print(f"adjoint vs central FD (64 coordinates, N = {n}): worst relative x {worst[0]:.3g} "
          f"v {worst[1]:.3g} m {worst[2]:.3g}")
    assert max(worst) <= 1e-4


def test_c3_adjoint_matches_oracle_fp64_adjoint_at_4096(hip):
    # The adjoint benchmark's problem (Plummer seed 1, eps 0.05, 200 KDK steps of 1e-3, checkpoints
    # every 20; target = the forward from positions perturbed by N(0, 0.01^2) and masses scaled by
    # (1 + 0.01 N(0, 1)); loss = sum |x_T - x_target|^2) at N = 4096, against the oracle's FP64 CPU
    # adjoint (every step kept, two unmerged half-kick VJPs per step from explicit Jacobian blocks).
    # Tolerances: mixed 1e-5 relative (x, v) and 1e-4 (m); all-FP64 forces 1e-8 (x, v) and 1e-6 (m).
    # The VJP reads each source's mass and cotangent as FP32 (its 48-byte source records) in both
    # precisions, so the mass gradient, which comes from the VJP alone, agrees to ~2e-8, and x and v,
    # where the force Jacobian is a small part of the gradient, to ~6e-10 (measured on an MI355X).
    # The CPU adjoint takes 5-30 s on the host's cores.
    n = 4096
    b = O.plummer(n, 1)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64)
    rng = np.random.default_rng(2026)
    xt, _ = O.sim_forward(x0 + 0.01 * rng.standard_normal(x0.shape), v0, m * (1 + 0.01 * rng.standard_normal(n)),
                          1.0, 0.05, 1e-3, 200)
    xo, vo = O.sim_forward(x0, v0, m, 1.0, 0.05, 1e-3, 200)
    ox, ov, om = O.sim_backward(2 * (xo - xt), np.zeros_like(vo))
    for prec, tol, tol_m in (("mixed", 1e-5, 1e-4), ("f64", 1e-8, 1e-6)):
        sim = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=200, checkpoint_every=20, precision=prec))
        xT, vT = sim.forward(x0, v0, m)
        assert rel(xT - x0, xo - x0) < tol, prec
        # the same cotangent on both sides (the oracle's final state), so only the adjoints differ
        gx, gv, gm = sim.backward(2 * (xo - xt), np.zeros_like(vT))
        print(f"c3 at N = {n}, {prec}: adjoint vs FP64 CPU adjoint x {rel(gx, ox):.3g} v {rel(gv, ov):.3g} "
              f"m {rel(gm, om):.3g}")
        assert rel(gx, ox) < tol and rel(gv, ov) < tol and rel(gm, om) < tol_m, prec


def test_checkpoint_interval_does_not_change_gradients(hip):
    x0, v0, m = plummer_state(1024, 6)
    out = []
    for K in (1, 7, 30):
        sim = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=30, checkpoint_every=K))
        xT, _ = sim.forward(x0, v0, m)
        out.append(sim.backward(2 * xT, np.ones_like(xT)))
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            assert np.array_equal(a, b)


def test_backward_before_forward_is_a_logic_error(hip):
    import lettuce_simulate_physics_amd as L

    fresh = L.HipBackend(0)
    try:
        with pytest.raises(L.LogicError):
            Simulation(fresh, SimParams()).backward(np.zeros((4, 3)), np.zeros((4, 3)))
        z = np.zeros((4, 3))
        rc = L.library().lhip_sim_backward(fresh._h, z.ctypes.data_as(ctypes.c_void_p),
                                           z.ctypes.data_as(ctypes.c_void_p), None, None, None)
        assert rc == L.LOGIC
    finally:
        fresh.close()


def test_backward_of_an_older_forward_is_a_logic_error(hip):
    # Each forward replaces the context's checkpoints: the backward of an earlier forward (another
    # Simulation, or an earlier graph of the autograd function) raises instead of returning the
    # later trajectory's gradients, also across methods (a direct forward after a tree forward).
    import lettuce_simulate_physics_amd as L

    x0, v0, m = plummer_state(256, 8)
    a = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=4, checkpoint_every=2))
    b = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=6, checkpoint_every=3))
    xa, _ = a.forward(x0, v0, m)
    ga = a.backward(xa, np.zeros_like(xa))
    xb, _ = b.forward(x0 + 0.01, v0, m)
    with pytest.raises(L.LogicError):
        a.backward(xa, np.zeros_like(xa))
    gb = b.backward(xb, np.zeros_like(xb))
    assert np.all(np.isfinite(gb[0]))
    xa2, _ = a.forward(x0, v0, m)
    assert np.array_equal(xa2, xa)
    for u, w in zip(a.backward(xa, np.zeros_like(xa)), ga):
        assert np.array_equal(u, w)
    tree = Simulation(hip, SimParams(method="tree", eps=0.05, dt=1e-3, steps=4, checkpoint_every=2))
    xt, _ = tree.forward(x0, v0, m)
    a.forward(x0, v0, m)
    with pytest.raises(L.LogicError):
        tree.backward(xt, np.zeros_like(xt))
    hip.set_precision("plain")

    torch = pytest.importorskip("torch")
    F = leapfrog_function()
    X1, X2 = (torch.tensor(x, requires_grad=True) for x in (x0, x0 + 0.01))
    V, M = torch.tensor(v0), torch.tensor(m)
    y1, _ = F.apply(X1, V, M, a)
    y2, _ = F.apply(X2, V, M, b)
    with pytest.raises(RuntimeError, match="another forward"):
        y1.sum().backward()
    y2.sum().backward()
    assert X2.grad is not None and torch.all(torch.isfinite(X2.grad))


def test_torch_autograd_function(hip):
    torch = pytest.importorskip("torch")
    x0, v0, m = plummer_state(512, 7)
    tgt = torch.from_numpy(x0 + 0.01)
    sim = Simulation(hip, SimParams(eps=0.05, dt=1e-3, steps=10, checkpoint_every=4))
    X, V, Mt = (torch.tensor(a, requires_grad=True) for a in (x0, v0, m))
    xT, vT = leapfrog_function().apply(X, V, Mt, sim)
    loss = ((xT - tgt) ** 2).sum()
    loss.backward()
    xT2, _ = sim.forward(x0, v0, m)
    gx, gv, gm = sim.backward(2 * (xT2 - tgt.numpy()), np.zeros_like(xT2))
    assert np.allclose(X.grad.numpy(), gx, rtol=1e-12, atol=0)
    assert np.allclose(V.grad.numpy(), gv, rtol=1e-12, atol=0)
    assert np.allclose(Mt.grad.numpy(), gm, rtol=1e-12, atol=0)
