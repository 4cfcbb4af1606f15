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
# Issue-slot ceiling of the pair loops. VALU instructions per interaction, read from the built
# kernels' gfx950 ISA (the 1-3 slot loops): 12 per direct pair, about 13.3 per walk body pair and
# 32.3 per walk cell term, each besides one v_rsq_f32. What an instruction slot and an rsq cost is
# measured live by the issue probe (lettuce_simulate_physics_amd/csrc/probe/issue_probe.hip): wave64
# streams of independent FMAs give the device's lane-instruction rate, and a stream of 12 FMAs plus
# one v_rsq_f32 per step against the 12 FMAs alone gives the cost of an rsq in FMA slots (0 when it
# issues beside the FMA stream).
VALU_DIRECT, VALU_P2P, VALU_P2C = 12.0, 13.3, 32.3
NOMINAL_LANE_RATE = 256 * 4 * 16 * 2.4e9  # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz

_probe = None


def issue_probe() -> dict:
    """The issue probe's measurements (run once per process)."""
    global _probe
    if _probe is None:
        import ctypes

        lib = ctypes.CDLL(os.path.join(ROOT, "lettuce_simulate_physics_amd", "libissue_probe.so"))
        lib.issue_probe_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        blocks, steps, chains = 256 * 8, 4096, 12
        ms = (ctypes.c_double * 3)()
        if lib.issue_probe_run(blocks, steps, ms) != 0:
            raise RuntimeError("issue probe failed")
        fma, rsq, mix = ms[0], ms[1], ms[2]
        lanes = blocks * 256 * steps * chains
        _probe = {"fma_ms": round(fma, 4), "rsq_ms": round(rsq, 4), "mix_ms": round(mix, 4),
                  "lane_rate_tps": round(lanes / (fma * 1e-3) / 1e12, 3),
                  "nominal_lane_rate_tps": round(NOMINAL_LANE_RATE / 1e12, 3),
                  "rsq_throughput_vs_fma": round(fma / rsq, 3),
                  "rsq_cost_fma_slots": round(max(0.0, chains * (mix - fma) / fma), 3),
                  "blocks": blocks, "steps": steps, "chains": chains}
    return _probe


def issue_sweep() -> dict:
    """Issue rate of single-form FP32 streams (the probe's second entry point) by instruction form,
    independent chains per lane and waves per SIMD (1024, 2048 and 4096 blocks of 4 waves)."""
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "lettuce_simulate_physics_amd", "libissue_probe.so"))
    lib.issue_probe_sweep.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_double)]
    forms = ["v_fmamk_f32", "v_fmac_f32", "v_fma_f32 (3 VGPR)", "v_mul_f32", "v_add_f32"]
    steps, rows = 4096, []
    for f, name in enumerate(forms):
        for chains in (8, 12, 24):
            for blocks in (1024, 2048, 4096):
                ms = ctypes.c_double()
                if lib.issue_probe_sweep(f, chains, blocks, steps, ctypes.byref(ms)) != 0:
                    raise RuntimeError("issue sweep failed")
                rate = blocks * 256 * steps * chains / (ms.value * 1e-3)
                rows.append({"form": name, "chains": chains, "waves_per_simd": blocks * 4 // 1024,
                             "ms": round(ms.value, 4), "lane_rate_tps": round(rate / 1e12, 2),
                             "x_nominal": round(rate / NOMINAL_LANE_RATE, 3)})
    return {"steps": steps, "nominal_lane_rate_tps": round(NOMINAL_LANE_RATE / 1e12, 3), "rows": rows}


def issue_ceiling(n_pp: float, n_pc: float, valu_pp: float, seconds: float) -> dict:
    """FP32 rate the instruction mix allows at the probe's issue rate, and the share of it achieved."""
    P = issue_probe()
    rsq = P["rsq_cost_fma_slots"]
    slots = (valu_pp + rsq) * n_pp + (VALU_P2C + rsq) * n_pc
    t_min = slots / (P["lane_rate_tps"] * 1e12)
    flops = FLOP_P2P * n_pp + FLOP_P2C * n_pc
    return {"valu_per_p2p": valu_pp, "valu_per_p2c": VALU_P2C, "rsq_cost_fma_slots": rsq,
            "lane_rate_tps": P["lane_rate_tps"], "ceiling": round(flops / t_min / 1e12, 2), "unit": "TFLOP/s",
            "frac": round(t_min / seconds, 4)}


def accum_probe(cases=((262144, 0.01), (65536, 0.05))) -> dict:
    """The accumulation probe (libaccum_probe.so): the mixed form's FP32 pair arithmetic summed in
    FP32, with Kahan compensation in FP32, in FP32 runs of 64 carried into FP64 (the library's
    choice) and term by term in FP64, against FP64 pair arithmetic, for every body of a Plummer
    sphere (a = 1, cut 10, seed 1) at each (n, eps); relative errors of |a| per body and the
    kernel time of each mode."""
    import ctypes

    import lettuce_simulate_physics_amd as L

    lib = ctypes.CDLL(os.path.join(ROOT, "lettuce_simulate_physics_amd", "libaccum_probe.so"))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.accum_probe_run.argtypes = [dp, dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
    names = ["fp64_reference", "fp32_sum", "fp32_kahan", "fp32_runs_of_64_fp64_carry", "fp64_sum_of_fp32_terms"]
    out = {}
    for n, eps in cases:
        x, _, m = L.plummer_f64(n, 1, 1.0, 10.0)
        cols = [np.ascontiguousarray(x[:, k]) for k in range(3)] + [np.ascontiguousarray(m)]
        res, ms = {}, {}
        for mode in range(5):
            a = np.zeros(3 * n)
            t = ctypes.c_double(0.0)
            rc = lib.accum_probe_run(*[c.ctypes.data_as(dp) for c in cols], n, eps, mode, a.ctypes.data_as(dp),
                                     ctypes.byref(t))
            if rc != 0:
                raise RuntimeError(f"accumulation probe failed (mode {mode}): {rc}")
            res[mode], ms[mode] = a.reshape(n, 3), t.value
        ref = res[0]
        mag = np.linalg.norm(ref, axis=1)
        case = {"n": n, "eps": eps, "pairs": n * n}
        for mode in range(1, 5):
            err = np.linalg.norm(res[mode] - ref, axis=1) / mag
            case[names[mode]] = {"median": float(np.median(err)), "p99": float(np.quantile(err, 0.99)),
                                 "max": float(err.max()), "ms": round(ms[mode], 4),
                                 "gpairs_per_s": round(n * n / (ms[mode] * 1e-3) / 1e9, 1)}
        case[names[0]] = {"ms": round(ms[0], 4)}
        out[str(n)] = case
    return out


def plummer(n: int, seed: int = 1):
    """The workload's initial conditions: the library's generator, bit-exact with the reference
    scenario 'plummer' (plummer.a = 1, plummer.cut = 10)."""
    import lettuce_simulate_physics_amd as L

    return L.plummer(n, seed, 1.0, 10.0)


def host_cpu() -> dict:
    """Model name, sockets and hardware threads of this host (lscpu, else /proc/cpuinfo)."""
    import subprocess

    info = {}
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name")
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    return {"cpu_model": model, "sockets": int(info["Socket(s)"]) if info.get("Socket(s)", "").isdigit() else None,
            "host_threads": os.cpu_count()}


def cpu_baseline(steps: int = 5):
    """Reference CPU backend (oracle/_ref, compiled from the reference sources) if present, else the
    oracle's restatement, on this host's cores with lettuce-bench's method (src/bench.cpp:169-198):
    one warm-up force call, then `steps` timed KDK steps at N = 2^20, OMP_PROC_BIND=close. Runs in a
    child process so the OpenMP settings reach the runtime before it starts."""
    import subprocess

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS", threads))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--n", str(N_BODIES), "--steps", str(steps),
                        "--theta", str(THETA), "--eps", str(EPS), "--dt", str(DT)],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline failed: {r.stderr[-2000:]}")
    c = json.loads(r.stdout.strip().splitlines()[-1])
    return dict(value=round(c["ginteractions_per_s"], 4), unit="Ginteractions/s", ms_per_step=round(c["ms_per_step"], 3),
                force_ms=round(c["force_ms"], 3), cores=c["threads"], kind=c["kind"], omp_proc_bind="close",
                **host_cpu(),
                sample=f"N=2^20 Plummer tree theta=0.6: 1 warm-up force call + {steps} KDK steps "
                       "(lettuce-bench method, src/bench.cpp:169-198)")


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


def run_steps(be, p, steps: int, dt: float, check=None):
    """`steps` KDK steps (lettuce-bench order). Without check: one wall-clock region around all the
    steps, bracketed by device synchronisations. With check(k) (run after step k, untimed): the
    steps are timed one by one, each bracketed the same way. Phase times come from times() after
    every step. Returns timing sums and counts."""
    import torch

    r = dict(wall=0.0, force=0.0, sort=0.0, build=0.0, walk=0.0, inter=0, p2p=0, p2c=0, p2m=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        be.kick(0.5 * dt)
        be.drift(dt)
        be.compute_forces(p, False)
        be.kick(0.5 * dt)
        t = be.times()  # resolves the kick/drift events (synchronises)
        for key, name in (("force", "force_ms"), ("sort", "sort_ms"), ("build", "build_ms"), ("walk", "walk_ms")):
            r[key] += t[name]
        r["inter"] += t["interactions"]
        r["p2p"] += t["p2p"]
        r["p2c"] += t["p2c"]
        r["p2m"] += t["p2c_mono"]
        if check:
            torch.cuda.synchronize()
            r["wall"] += time.perf_counter() - t0
            check(k)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
    torch.cuda.synchronize()
    if not check:
        r["wall"] += time.perf_counter() - t0
    return r


def force_errors(be, eps: float, sample: int, seed: int):
    """Per-body relative force error of the last force call against the device FP64 direct sum on a
    seeded sample of bodies: (median, p99)."""
    n = be.size()
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, min(sample, n), replace=False)).astype(np.uint32)
    mode, _ = be.precision()
    if mode == "mixed":
        ax, ay, az, _ = be.accelerations_f64()
    else:
        ax, ay, az, _ = [a.astype(np.float64) for a in be.accelerations()]
    a = np.stack([ax[idx], ay[idx], az[idx]], 1)
    ref, _ = be.direct_f64_targets(idx, 1.0, eps)
    e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.median(e)), float(np.quantile(e, 0.99))


def energy(be, p):
    be.compute_forces(p, True)
    d = be.diagnostics()
    return d["ke"] + d["pe"]


def summary(r, steps):
    return {"ms_per_step": round(r["wall"] * 1e3 / steps, 4), "force_ms": round(r["force"] / steps, 4),
            "sort_ms": round(r["sort"] / steps, 4), "build_ms": round(r["build"] / steps, 4),
            "walk_ms": round(r["walk"] / steps, 4),
            "ginteractions_per_s": round(r["inter"] / (r["force"] * 1e-3) / 1e9, 3),
            "interactions_per_body_step": None}


def mixed_leg(be, b, steps: int, warmup: int, wopt: dict, plain_force_ms: float):
    """The same 2^20-body Plummer tree step in the mixed-precision form (FP64 state, FP32
    relative-coordinate pair arithmetic, FP64 carry, and the mixed form's walk options: by default
    the error-budgeted far field). Force errors of both forms at the initial state against the FP64
    direct sum on 4096 seeded bodies."""
    import lettuce_simulate_physics_amd as L

    p = L.ForceParams(G=1.0, eps=EPS, theta=THETA, method="tree")
    be.set_precision("plain")
    be.upload(b)
    be.compute_forces(p, False)
    pmed, pp99 = force_errors(be, EPS, 4096, 11)
    be.set_precision("mixed", **wopt)
    be.upload(b)
    be.compute_forces(p, False)
    med, p99 = force_errors(be, EPS, 4096, 11)
    run_steps(be, p, warmup, DT)
    r = run_steps(be, p, steps, DT)
    out = summary(r, steps)
    out["interactions_per_body_step"] = round(r["inter"] / steps / len(b), 2)
    out["monopole_fraction_of_p2c"] = round(r["p2m"] / max(r["p2c"], 1), 4)
    opts = be.precision()[1]
    be.set_precision("plain")
    return dict(workload="plummer_tree_1M_mixed", walk_options=opts, **out,
                force_speedup_vs_plain=round(plain_force_ms / out["force_ms"], 3),
                median_force_error=med, p99_force_error=p99, plain_median_force_error=pmed,
                plain_p99_force_error=pp99,
                error_reference="FP64 direct sum on 4096 seeded bodies at the initial state")


def disk_leg(be, nstars: int, steps: int, wopt: dict):
    """Rotating sech^2 disk, nstars equal-mass stars (total mass 1, Rd 1, cut 10 Rd, z0 0.1,
    dispersion 0.1 vc) around a point mass 10; tree theta 0.5, eps 0.02, KDK dt 5e-4. Runs the plain
    and the mixed form from the same initial state; force error vs FP64 direct on 8192 seeded bodies
    every 10 steps, energy drift over the run (tree potential)."""
    import lettuce_simulate_physics_amd as L

    b = L.disk_sech2(nstars, seed=1, core=10.0, mass=1.0, rd=1.0, rmax=10.0, z0=0.1, sigma=0.1)
    p = L.ForceParams(G=1.0, eps=0.02, theta=0.5, method="tree")
    res = {}
    for mode, o in (("plain", {}), ("mixed", wopt)):
        be.set_precision(mode, **o)
        be.upload(b)
        e0 = energy(be, p)
        errs = [force_errors(be, 0.02, 8192, 100)]

        def check(k):
            if k % 10 == 0:
                errs.append(force_errors(be, 0.02, 8192, 100 + k))

        r = run_steps(be, p, steps, 5e-4, check)
        e1 = energy(be, p)
        out = summary(r, steps)
        out["interactions_per_body_step"] = round(r["inter"] / steps / len(b), 2)
        out["walk_options"] = be.precision()[1]
        out["monopole_fraction_of_p2c"] = round(r["p2m"] / max(r["p2c"], 1), 4)
        out["force_error_median_per_check"] = [float(f"{m:.4g}") for m, _ in errs]
        out["force_error_p99_per_check"] = [float(f"{q:.4g}") for _, q in errs]
        out["energy_drift"] = float(f"{abs(e1 - e0) / abs(e0):.4g}")
        res[mode] = out
    be.set_precision("plain")
    return {"workload": f"disk_sech2_{nstars}_plus_core", "steps": steps, "theta": 0.5, "eps": 0.02, "dt": 5e-4,
            "checks_every": 10, **res,
            "mixed_force_speedup_vs_plain": round(res["plain"]["force_ms"] / res["mixed"]["force_ms"], 3)}


def tree_adjoint_leg(be, nstars: int, steps: int):
    """Adjoint through the frozen-topology tree walk: two sech^2 disks of nstars stars and a core of 10
    each, inclined 30 and 60 degrees, 20 Rd apart on a parabolic approach with pericentre 4 Rd; tree
    theta 0.6 (mixed form, its default walk options), eps 0.02, KDK dt 1e-3, checkpoints every 20
    steps. Loss: squared distance of disk 2's final centre of mass from a seeded target point.
    Times one forward and one backward (device events); gradients in (pericentre, incl1, incl2).
    Also with a checkpoint every step (no recompute). Cross-check: the same problem at 2 x 16384 stars
    through the tree (the mixed far field, and every quadrupole kept) and through the direct-sum
    adjoint, the largest gradient difference relative to the largest component, and dL/d(incl1)
    against central differences of the frozen-topology forward (replays, h = 1e-2 and 5e-3,
    Richardson)."""
    import lettuce_simulate_physics_amd as L
    from lettuce_simulate_physics_amd.autograd import SimParams, Simulation

    target = np.random.default_rng(2027).standard_normal(3)

    def enc(ns, **kw):
        return L.Collision(**{**dict(nstars=ns, seed=1, pericentre=4.0, incl1=np.pi / 6, incl2=np.pi / 3), **kw})

    def com2(ns, xT, m):
        w = np.zeros(len(m))
        w[ns + 1:] = m[ns + 1:] / m[ns + 1:].sum()
        return w, (w[:, None] * xT).sum(0)

    def run(ns, method, k, every=20, walk=None, tree_mode="build"):
        C = enc(ns)
        x0, v0, m = C.generate()
        sim = Simulation(be, SimParams(method=method, theta=0.6, eps=0.02, dt=1e-3, steps=k, checkpoint_every=every,
                                       walk=walk, tree_mode=tree_mode))
        xT, vT = sim.forward(x0, v0, m)
        fwd_ms, _ = sim.times()
        w, com = com2(ns, xT, m)
        gx, gv, _ = sim.backward(2 * w[:, None] * (com - target)[None, :], np.zeros_like(vT))
        _, bwd_ms = sim.times()
        return fwd_ms, bwd_ms, float(((com - target) ** 2).sum()), C.vjp(gx, gv)

    fwd_ms, bwd_ms, loss, g = run(nstars, "tree", steps)
    # the same with a checkpoint every step (no recompute: HBM for time)
    fwd1, bwd1, _, g1 = run(nstars, "tree", steps, 1)
    _, _, _, gd = run(16384, "direct", 100)
    _, _, _, gq = run(16384, "tree", 100, walk=dict(group_max=128, small_leaf=4, quad_tau=0.0))
    _, _, _, gt = run(16384, "tree", 100, tree_mode="record")
    rep = Simulation(be, SimParams(method="tree", theta=0.6, eps=0.02, dt=1e-3, steps=100, checkpoint_every=20,
                                   tree_mode="replay"))
    est = []
    for h in (1e-2, 5e-3):
        vals = []
        for sgn in (1, -1):
            xp, vp, mp = enc(16384, incl1=np.pi / 6 + sgn * h).generate()
            xq, _ = rep.forward(xp, vp, mp)
            vals.append(float(((com2(16384, xq, mp)[1] - target) ** 2).sum()))
        est.append((vals[0] - vals[1]) / (2 * h))
    fd_incl1 = (4 * est[1] - est[0]) / 3
    be.set_precision("plain")
    return {"workload": f"two_disk_encounter_2x{nstars}_tree_adjoint", "steps": steps, "checkpoint_every": 20,
            "forward_ms_per_step": round(fwd_ms / steps, 4), "adjoint_ms_per_step": round(bwd_ms / steps, 4),
            "adjoint_forward_ratio": round(bwd_ms / fwd_ms, 3), "loss": loss,
            "grad_pericentre": float(g[0]), "grad_incl1": float(g[1]), "grad_incl2": float(g[2]),
            "checkpoint_every_step": {"forward_ms_per_step": round(fwd1 / steps, 4),
                                      "adjoint_ms_per_step": round(bwd1 / steps, 4),
                                      "adjoint_forward_ratio": round(bwd1 / fwd1, 3),
                                      "max_grad_diff": float(np.abs(g1 - g).max())},
            "cross_check_2x16384_100_steps": {
                "tree": [float(v) for v in gt], "direct": [float(v) for v in gd],
                "max_diff_over_largest": float(np.abs(gt - gd).max() / np.abs(gd).max()),
                "tree_every_quadrupole": [float(v) for v in gq],
                "incl1_frozen_forward_fd": float(fd_incl1),
                "incl1_adjoint_vs_fd": float(abs(gt[1] - fd_incl1) / abs(fd_incl1)),
                "incl1_note": "the mixed far field evaluates cells with small quadrupoles as point masses, "
                              "which drops the quadrupoles that carry disk 1's orientation: dL/d(incl1) "
                              "matches the frozen forward's differences but not the direct sum's; with "
                              "every quadrupole kept it is within 3.2% of the direct sum's"}}


# Algorithmic HBM bytes per body of the force step's memory-bound phases (the walk is VALU-bound):
# the sort phase is the bounding box (12), keys (12 in + 12 out), one pass of (key, index) in and
# out (24, the single-pass lower bound) and the permute (32 + 32 + 4); the build reads the keys (8)
# and posm (16) and writes each node's 32 + 64 B records and reads the 64-B moments of each child
# once (96 + 64 per node).
SORT_BYTES_PER_BODY = 12 + 24 + 24 + 68
BUILD_BYTES_PER_BODY = 8 + 16
BUILD_BYTES_PER_NODE = 96 + 64
PEAK_HBM = 8.0e12


def hbm_phases(r, steps: int, n: int, nodes: int):
    """Achieved HBM rates of the sort and build phases from their event-timed averages."""
    out = {}
    for name, ms, nbytes in (("sort", r["sort"] / steps, SORT_BYTES_PER_BODY * n),
                             ("build", r["build"] / steps, BUILD_BYTES_PER_BODY * n + BUILD_BYTES_PER_NODE * nodes)):
        rate = nbytes / (ms * 1e-3)
        out[name] = {"ms": round(ms, 4), "algorithmic_bytes": int(nbytes), "GB_s": round(rate / 1e9, 1),
                     "frac": round(rate / PEAK_HBM, 4)}
    out["nodes"] = nodes
    return out


def direct_leg(sizes=(16384, 65536), reps: int = 5):
    """The direct sum (src/cpu_backend.cpp:353-392) at the 16,384- and 65,536-body Plummer sizes:
    device time of the force call, pair rate and FP32 rate (20 flop per pair) against the peak."""
    import lettuce_simulate_physics_amd as L

    be = L.HipBackend(0)
    out = {}
    p = L.ForceParams(G=1.0, eps=EPS, theta=THETA, method="direct")
    for n in sizes:
        be.upload(plummer(n))
        be.compute_forces(p, False)
        ms = []
        for _ in range(reps):
            be.compute_forces(p, False)
            ms.append(be.times()["walk_ms"])
        t = float(np.median(ms))
        pairs = n * (n - 1)
        out[str(n)] = {"ms": round(t, 4), "gpairs_per_s": round(pairs / (t * 1e-3) / 1e9, 1),
                       "tflops": round(FLOP_P2P * pairs / (t * 1e-3) / 1e12, 2),
                       "frac": round(FLOP_P2P * pairs / (t * 1e-3) / PEAK_FP32, 4),
                       "issue_ceiling": issue_ceiling(pairs, 0, VALU_DIRECT, t * 1e-3)}
    be.close()
    return out


def let_leg(b, reps: int = 5):
    """Two contexts on this GPU as two ranks of the reference's distributed tree step
    (src/parallel.cpp:260-295): the workload's sphere split at x = 0. Each rank builds its tree,
    exports its LET for the other's box and computes its forces with the other's LET. Medians over
    `reps` (after one untimed round) of the export (host clock: device passes, two small read-backs
    and the copy out) and of the walk with and without the remote tree (device time)."""
    import lettuce_simulate_physics_amd as L

    p = L.ForceParams(G=1.0, eps=EPS, theta=THETA, method="tree")
    halves = [b.x < 0, b.x >= 0]
    r = [L.HipBackend(0), L.HipBackend(0)]
    for k in range(2):
        r[k].upload(L.Bodies(*[np.ascontiguousarray(a[halves[k]]) for a in b.arrays()]))
    rows = []
    for rep in range(reps + 1):
        row = {}
        for k in range(2):
            r[k].build_tree(p)
        boxes = [r[k].local_bounds() for k in range(2)]
        t0 = time.perf_counter()
        lets = [r[1].export_let(boxes[0], p), r[0].export_let(boxes[1], p)]
        row["export_ms"] = (time.perf_counter() - t0) * 1e3 / 2
        local_ms, total_ms, inter_local, inter_remote = 0.0, 0.0, 0, 0
        for k in range(2):
            r[k].compute_forces_with(p, False, [])
            t = r[k].times()
            local_ms += t["walk_ms"]
            inter_local += t["interactions"]
            r[k].compute_forces_with(p, False, [lets[k]])
            t = r[k].times()
            total_ms += t["walk_ms"]
            inter_remote += t["interactions"]
        row.update(walk_local_ms=local_ms / 2, walk_with_let_ms=total_ms / 2,
                   remote_inter=inter_remote - inter_local, nodes=[len(t_.nodes) for t_ in lets],
                   bodies=[len(t_.bx) for t_ in lets])
        if rep > 0:
            rows.append(row)
    med = lambda k: float(np.median([x[k] for x in rows]))  # noqa: E731
    remote_ms = med("walk_with_let_ms") - med("walk_local_ms")
    out = {"workload": "plummer_1M_two_ranks_x_split", "ranks_bodies": [int(h.sum()) for h in halves],
           "let_nodes": rows[0]["nodes"], "let_bodies": rows[0]["bodies"],
           "let_bytes": [60 * a + 16 * c for a, c in zip(rows[0]["nodes"], rows[0]["bodies"])],
           "export_ms": round(med("export_ms"), 4), "walk_local_ms": round(med("walk_local_ms"), 4),
           "walk_with_let_ms": round(med("walk_with_let_ms"), 4), "remote_pass_ms": round(remote_ms, 4),
           "remote_interactions": int(rows[0]["remote_inter"]),
           "remote_ginteractions_per_s": round(rows[0]["remote_inter"] / 2 / max(remote_ms, 1e-9) / 1e6, 2)}
    for be in r:
        be.close()
    return out


# The reference's published CUDA numbers (docs/running.md:71-81; H200 SXM, lettuce-bench method=tree
# n=1e6,1e7,1e8 steps=3, Hénon-unit Plummer a = 3 pi / 16, cut 20, seed 1, eps 0.01, theta 0.6,
# dt 1e-3): force ms per step and Ginteractions/s.
PUBLISHED_TREE = {1000000: (5.40, 472.0), 10000000: (50.9, 532.0), 100000000: (532.0, 538.0)}


def henon_leg(sizes, steps: int = 3):
    """The reference's published benchmark configuration, like for like: lettuce-bench's default
    Plummer scenario (Hénon units: a = 3 pi / 16, cut 20, seed 1, masses 1/N), tree theta 0.6, eps
    0.01, dt 1e-3; one warm-up force call, then `steps` timed KDK steps (src/bench.cpp:169-198).
    Force ms and Ginteractions/s beside the published H200 figures; bodies resident on the device
    before timing (the PCIe upload is reported apart)."""
    import lettuce_simulate_physics_amd as L

    be = L.HipBackend(0)  # its own context: the 1e8 buffers go with it
    p = L.ForceParams(G=1.0, eps=EPS, theta=THETA, method="tree")
    rows = {}
    for n in sizes:
        t0 = time.perf_counter()
        b = L.plummer(n, 1, 3 * np.pi / 16, 20.0)
        gen_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        be.upload(b)
        upload_s = time.perf_counter() - t0
        del b
        be.compute_forces(p, False)
        r = run_steps(be, p, steps, DT)
        force_ms = r["force"] / steps
        ginter = r["inter"] / (r["force"] * 1e-3) / 1e9
        row = {"n": n, "steps": steps, "ms_per_step": round(r["wall"] * 1e3 / steps, 3), "force_ms": round(force_ms, 3),
               "sort_ms": round(r["sort"] / steps, 3), "build_ms": round(r["build"] / steps, 3),
               "walk_ms": round(r["walk"] / steps, 3), "ginteractions_per_s": round(ginter, 1),
               "interactions_per_body": round(r["inter"] / steps / n, 1),
               "upload_s": round(upload_s, 3), "generate_s": round(gen_s, 1)}
        if n in PUBLISHED_TREE:
            ms, gi = PUBLISHED_TREE[n]
            row.update(published_h200_force_ms=ms, published_h200_ginteractions_per_s=gi,
                       force_time_speedup_vs_h200=round(ms / force_ms, 3))
        rows[str(n)] = row
        print(f"henon n={n}: force {force_ms:.3f} ms, {ginter:.1f} Gint/s", file=sys.stderr, flush=True)
    be.close()
    return {"workload": "henon_plummer_tree", "plummer_a": round(3 * np.pi / 16, 6), "plummer_cut": 20.0, "seed": 1,
            "theta": THETA, "eps": EPS, "dt": DT, "published": "docs/running.md:71-81 (H200 SXM, CUDA backend)",
            "sizes": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_BODIES)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--mixed-group", type=int, default=None, help="walk group size of the mixed form")
    ap.add_argument("--mixed-small-leaf", type=int, default=None)
    ap.add_argument("--mixed-tau", type=float, default=None, help="far-field quadrupole threshold (0: all kept)")
    ap.add_argument("--disk-stars", type=int, default=4194304)
    ap.add_argument("--disk-steps", type=int, default=50)
    ap.add_argument("--encounter-stars", type=int, default=524288, help="stars per disk of the tree adjoint leg")
    ap.add_argument("--encounter-steps", type=int, default=100, help="0 skips the tree adjoint leg")
    ap.add_argument("--no-let", action="store_true", help="skip the two-rank LET leg")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-sum leg")
    ap.add_argument("--henon-sizes", type=str, default="1e6,1e7,1e8",
                    help="sizes of the like-for-like leg against the published numbers ('' skips it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy 3 times (HBM counter calibration)")
    ap.add_argument("--issue-probe", action="store_true", help="only run the issue-rate probe and print it")
    ap.add_argument("--accum-probe", action="store_true",
                    help="only run the accumulation probe (FP32, Kahan, runs with an FP64 carry) and print it")
    ap.add_argument("--issue-sweep", action="store_true",
                    help="only run the issue-rate sweep (instruction form x chains x waves per SIMD) and print it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.issue_probe:
        print(json.dumps(issue_probe()), flush=True)
        return
    if args.accum_probe:
        print(json.dumps(accum_probe()), flush=True)
        return
    if args.issue_sweep:
        print(json.dumps(issue_sweep()), flush=True)
        return

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
    run_steps(be, p, args.warmup, DT)
    torch.cuda.synchronize()
    k = args.steps
    r = run_steps(be, p, k, DT)
    wall_ms = r["wall"] * 1e3 / k
    force_ms, walk_ms = r["force"] / k, r["walk"] / k
    ginter = r["inter"] / (r["force"] * 1e-3) / 1e9
    walk_flops = (FLOP_P2P * r["p2p"] + FLOP_P2C * r["p2c"]) / k
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
        "sort_ms": round(r["sort"] / k, 4),
        "build_ms": round(r["build"] / k, 4),
        "walk_ms": round(walk_ms, 4),
        "interactions_per_body": round(r["inter"] / k / args.n, 2),
        "roofline": {"kernel": "walk_kernel", "bound": "valu", "achieved": round(achieved / 1e12, 3),
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32, 4),
                     "traffic": traffic,
                     "issue_ceiling": issue_ceiling(r["p2p"] / k, r["p2c"] / k, VALU_P2P, walk_ms * 1e-3)},
        "hbm_phases": hbm_phases(r, k, args.n, be.tree_counts()[0]),
        "issue_probe": issue_probe(),
        "cpu_baseline": None,
        "mixed": None,
        "disk": None,
        "adjoint": None,
    }
    if not args.no_mixed:
        wopt = dict(group_max=args.mixed_group, small_leaf=args.mixed_small_leaf, quad_tau=args.mixed_tau)
        out["mixed"] = mixed_leg(be, b, k, args.warmup, wopt, force_ms)
        if args.disk_steps > 0:
            out["disk"] = disk_leg(be, args.disk_stars, args.disk_steps, wopt)
    if not args.no_direct:
        out["direct"] = direct_leg()
    if not args.no_let:
        out["let"] = let_leg(b)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(be)
        out["adjoint_forward_ratio"] = out["adjoint"]["adjoint_forward_ratio"]
        if args.encounter_steps > 0:
            out["tree_adjoint"] = tree_adjoint_leg(be, args.encounter_stars, args.encounter_steps)
    if args.henon_sizes:
        out["henon"] = henon_leg([int(float(v)) for v in args.henon_sizes.split(",")])
        h1 = out["henon"]["sizes"].get("1000000")
        if h1:
            # value's metric on the published configuration (the Hénon-unit 1e6 sphere) over the
            # published H200 figure for it
            out["vs_baseline"] = round(h1["ginteractions_per_s"] / PUBLISHED_TREE[1000000][1], 3)
            out["vs_baseline_basis"] = ("Ginteractions/s of the henon leg at n=1e6 (a=3pi/16, cut 20) over the "
                                        "published H200 472 Gint/s (docs/running.md:78); force ms "
                                        f"{h1['force_ms']} vs 5.40")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)
    be.close()


if __name__ == "__main__":
    main()
