"""CPU baseline timing: test infrastructure, run only by bench.py's cpu_baseline leg, in a child
process of its own so that OMP_NUM_THREADS and OMP_PROC_BIND reach the OpenMP runtime before it
starts.

Times the reference's own CPU backend (oracle/_ref/libref.so, compiled from the reference sources)
where it was built, else this directory's restatement of it, with lettuce-bench's method
(src/bench.cpp:46,169-198): one warm-up force call, then K timed KDK steps (kick dt/2, drift dt,
compute_forces, kick dt/2); ms/step is the wall time per step and Ginteractions/s the interactions
over the force time. Workload: the Plummer sphere (a = 1, cut 10, seed 1, the reference generator's
output bit for bit), tree theta 0.6, eps 0.01, dt 1e-3. Prints one JSON object.

    OMP_PROC_BIND=close OMP_NUM_THREADS=16 python -m oracle.cpu_bench --n 1048576 --steps 5
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--eps", type=float, default=0.01)
    ap.add_argument("--dt", type=float, default=1e-3)
    args = ap.parse_args()

    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    if O.ref_available():
        be, kind = O.RefCpu(threads), "reference"
    else:
        be, kind = O.Oracle(threads), "port"
    b = O.plummer(args.n, 1, 1.0, 10.0)
    be.upload(b)
    be.compute_forces(1.0, args.eps, args.theta, 1, False)  # the warm-up force call
    force_s, inter, wall = 0.0, 0, 0.0
    for _ in range(args.steps):
        t0 = time.perf_counter()
        be.kick(0.5 * args.dt)
        be.drift(args.dt)
        tf = time.perf_counter()
        be.compute_forces(1.0, args.eps, args.theta, 1, False)
        tf = time.perf_counter() - tf
        be.kick(0.5 * args.dt)
        wall += time.perf_counter() - t0
        if kind == "reference":
            t = be.times()
            force_s += t["force_ms"] * 1e-3  # the backend's own force time, as lettuce-bench reads it
            inter += t["interactions"] - args.n  # its count includes each body's own entry
        else:
            force_s += tf
            inter += be.counts()["interactions"]
    wall /= args.steps
    print(json.dumps({"kind": kind, "threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
                      "ms_per_step": wall * 1e3, "force_ms": force_s * 1e3 / args.steps,
                      "ginteractions_per_s": inter / force_s / 1e9, "steps": args.steps, "n": args.n}))


if __name__ == "__main__":
    main()
