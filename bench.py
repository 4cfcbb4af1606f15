"""Benchmark: the 2^20-body Plummer tree step on one MI355X (lettuce-bench semantics).

Workload: Plummer sphere, N = 1,048,576 equal masses (G = 1, M = 1, a = 1, cut at 10 a, seed 1),
Barnes-Hut with quadrupoles, theta = 0.6, leaf <= 16, eps = 0.01, KDK leapfrog dt = 1e-3. A step is
kick(dt/2) + drift(dt) + compute_forces (Morton keys, sort, octree, multipoles, walk) + kick(dt/2),
as in the reference's lettuce-bench (src/bench.cpp:169-198): W untimed warm-up steps after one
initial force call, then K timed steps. Ginteractions/s = interactions / force time, the reference's
definition; ms/step is wall time per step. Bodies are resident on the device before timing starts.

Prints one JSON line. Also reports the dominant kernel's roofline (the walk: FP32 VALU bound, flops
counted as 20 per body-body and 51 per body-node interaction) and the reference CPU path timed on
this host's cores on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BODIES = 1 << 20
THETA, EPS, DT = 0.6, 0.01, 1e-3
FLOP_P2P, FLOP_P2C = 20, 51
PEAK_FP32 = 157.3e12  # MI355X published FP32 vector peak (packed issue)


def plummer(n: int, seed: int = 1):
    """The workload's initial conditions (same generator as the reference scenario 'plummer')."""
    import oracle as O  # the generator only; never the measured path

    return O.plummer(n, seed, 1.0, 10.0)


def cpu_baseline(b, steps: int = 2):
    """Reference CPU backend (oracle/_ref, compiled from the reference sources) if present, else the
    oracle's restatement, on this host's cores: 1 warm-up force call + `steps` KDK steps at N = 2^20."""
    import oracle as O

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    if O.ref_available():
        be, kind = O.RefCpu(threads), "reference"
    else:
        be, kind = O.Oracle(threads), "port"
    be.upload(b)
    be.compute_forces(1.0, EPS, THETA, 1, False)
    t0 = time.perf_counter()
    for _ in range(steps):
        be.kick(0.5 * DT)
        be.drift(DT)
        be.compute_forces(1.0, EPS, THETA, 1, False)
        be.kick(0.5 * DT)
    wall = (time.perf_counter() - t0) / steps
    if kind == "reference":
        inter = be.times()["interactions"]  # the reference CPU count includes each body's self entry
        inter -= len(b)
    else:
        inter = be.counts()["interactions"]
    return dict(value=inter / wall / 1e9, unit="Ginteractions/s", ms_per_step=wall * 1e3, cores=threads, kind=kind,
                sample=f"N=2^20 Plummer tree theta=0.6 step: 1 warm-up force call + {steps} KDK steps")


def adjoint_leg(be):
    """Adjoint through direct summation: Plummer N = 65,536 (a = 1, cut 10), eps 0.05, 200 KDK steps
    of dt 1e-3, checkpoints every 20 steps with recompute. Target snapshot: the forward run from
    positions perturbed by N(0, 0.01^2) and masses scaled by (1 + 0.01 N(0, 1)); loss = sum of squared
    final-position differences. Times one forward and one backward (device events)."""
    from lettuce_simulate_physics_amd.autograd import SimParams, Simulation

    n, steps = 65536, 200
    b = plummer(n)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64)
    rng = np.random.default_rng(2026)
    sim = Simulation(be, SimParams(G=1.0, eps=0.05, dt=1e-3, steps=steps, checkpoint_every=20))
    xt, _ = sim.forward(x0 + 0.01 * rng.standard_normal(x0.shape), v0, m * (1 + 0.01 * rng.standard_normal(n)))
    xT, vT = sim.forward(x0, v0, m)
    fwd_ms, _ = sim.times()
    gx, gv, gm = sim.backward(2 * (xT - xt), np.zeros_like(vT))
    _, bwd_ms = sim.times()
    assert np.all(np.isfinite(gx)) and np.all(np.isfinite(gm))
    pairs = float(n) * (n - 1)
    return {"workload": "plummer_direct_adjoint_65536", "steps": steps, "checkpoint_every": 20,
            "forward_ms_per_step": round(fwd_ms / steps, 4), "adjoint_ms_per_step": round(bwd_ms / steps, 4),
            "adjoint_forward_ratio": round(bwd_ms / fwd_ms, 3),
            "forward_gpairs_per_s": round(pairs * (steps + 1) / (fwd_ms * 1e-3) / 1e9, 1),
            "loss": float(((xT - xt) ** 2).sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4, "launches": 3}), flush=True)
        return

    import lettuce_simulate_physics_amd as L

    be = L.HipBackend(0)
    b = plummer(args.n)
    p = L.ForceParams(G=1.0, eps=EPS, theta=THETA, method="tree")
    be.upload(b)
    be.compute_forces(p, False)
    for _ in range(args.warmup):
        be.kick(0.5 * DT)
        be.drift(DT)
        be.compute_forces(p, False)
        be.kick(0.5 * DT)
    be.times()
    torch.cuda.synchronize()

    force = sort = build = walk = 0.0
    inter = p2p = p2c = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        be.kick(0.5 * DT)
        be.drift(DT)
        be.compute_forces(p, False)
        be.kick(0.5 * DT)
        t = be.times()  # resolves the kick/drift events (synchronises)
        force += t["force_ms"]
        sort += t["sort_ms"]
        build += t["build_ms"]
        walk += t["walk_ms"]
        inter += t["interactions"]
        p2p += t["p2p"]
        p2c += t["p2c"]
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    k = args.steps
    force_ms, walk_ms = force / k, walk / k
    ginter = inter / (force * 1e-3) / 1e9
    walk_flops = (FLOP_P2P * p2p + FLOP_P2C * p2c) / k
    achieved = walk_flops / (walk_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "walk_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    out = {
        "metric": "tree ms/step & Ginteractions/s at N=1M theta=0.6; adjoint/forward step-time ratio",
        "value": round(ginter, 3),
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plummer_tree_1M", "n": args.n, "theta": THETA, "eps": EPS, "dt": DT,
                   "leaf_max": 16, "group_max": 128, "plummer_a": 1.0, "plummer_cut": 10.0, "seed": 1},
        "force_ms": round(force_ms, 4),
        "sort_ms": round(sort / k, 4),
        "build_ms": round(build / k, 4),
        "walk_ms": round(walk_ms, 4),
        "interactions_per_body": round(inter / k / args.n, 2),
        "roofline": {"kernel": "walk_kernel", "bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32, 4),
                     "traffic": traffic},
        "cpu_baseline": None,
        "adjoint": None,
    }
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(be)
        out["adjoint_forward_ratio"] = out["adjoint"]["adjoint_forward_ratio"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(b, args.cpu_steps)
        out["cpu_baseline"]["value"] = round(out["cpu_baseline"]["value"], 4)
    print(json.dumps(out), flush=True)
    be.close()


if __name__ == "__main__":
    main()
