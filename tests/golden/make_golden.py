"""Writes the golden fixtures under tests/golden/ from the reference's own CPU backend and scenario
generator (oracle/_ref/libref.so, compiled from the reference sources by oracle/Makefile).

Run in the container that has the reference sources:  python tests/golden/make_golden.py
(``--let`` writes only the LET fixture, let.npz).
The fixtures are data only: inputs (as generator parameters plus digests of the generated arrays)
and the reference's outputs (accelerations of sampled bodies, energies, interaction counts, digests
of the sort permutation, keys and canonical octree).
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from helpers import canon_from_ref_export, digest, morton_keys_np  # noqa: E402

PLUMMER_CASES = [  # (n, seed, a, cut)
    (16384, 1, 1.0, 10.0),
    (1048576, 1, 1.0, 10.0),
    (65536, 1, 1.0, 10.0),
    (10000, 1, 3 * math.pi / 16, 20.0),
    (4096, 7, 1.0, 10.0),
]


def ref_plummer(n, seed, a, cut):
    return O.ref_scenario("plummer", n, seed, f"plummer.a={a!r};plummer.cut={cut!r}")


def cube_bodies(n: int, seed: int) -> O.Bodies:
    """Uniform cube [-1, 1]^3, velocities in [-1, 1], masses U(0.5, 1.5)/n (numpy generator)."""
    rng = np.random.default_rng(seed)
    b = O.Bodies.empty(n)
    for a in (b.x, b.y, b.z, b.vx, b.vy, b.vz):
        a[:] = rng.uniform(-1, 1, n).astype(np.float32)
    b.m[:] = (rng.uniform(0.5, 1.5, n) / n).astype(np.float32)
    b.id[:] = np.arange(n, dtype=np.uint32)
    return b


def coincident_set() -> O.Bodies:
    """300 cube bodies; 150 at one point (a run of identical keys) and a clump of 40 within 1e-7."""
    b = cube_bodies(300, 7)
    b.x[:150], b.y[:150], b.z[:150] = 0.25, -0.5, 0.125
    k = np.arange(40, dtype=np.float32)
    b.x[150:190] = np.float32(-0.3) + np.float32(1e-7) * k
    b.y[150:190], b.z[150:190] = 0.1, 0.2
    return b


def tree_record(R: O.RefCpu, b: O.Bodies, eps: float, theta: float, potential: bool) -> dict:
    """Reference tree results for bodies b: permutation, keys, canonical octree, accelerations."""
    acc = R.accel_by_id(b, 1.0, eps, theta, 1, potential)
    t = R.times()
    o = R.download()
    costs = R.body_costs()
    keys = morton_keys_np(o.x, o.y, o.z)
    tree = R.export_tree(1.0, eps, theta)
    canon = canon_from_ref_export(tree)
    return dict(acc=acc, interactions=int(t["interactions"]), ids=o.id.copy(), keys=keys, canon=canon, costs=costs)


def main():
    R = O.RefCpu()
    out = {}
    # 1. Generator digests.
    gen = []
    for n, seed, a, cut in PLUMMER_CASES:
        b = ref_plummer(n, seed, a, cut)
        gen.append(dict(n=n, seed=seed, a=a, cut=cut, digest=digest(*b.arrays()),
                        head=[float(b.x[0]), float(b.y[0]), float(b.z[0]), float(b.vx[0])]))
    out["plummer"] = gen

    # 2. Direct summation, N = 16384 Plummer (a = 1, cut 10), eps 0.01, 10 KDK steps of dt 1e-3.
    b = ref_plummer(16384, 1, 1.0, 10.0)
    eps, dt = 0.01, 1e-3
    rng = np.random.default_rng(2024)
    sample = np.sort(rng.choice(16384, 1024, replace=False)).astype(np.int64)
    acc = R.accel_by_id(b, 1.0, eps, 0.6, 0)
    R.upload(b)
    R.compute_forces(1.0, eps, 0.6, 0, True)
    d0 = R.diagnostics()
    exact = R.exact_pe(1.0, eps)
    series = [[d0[k] for k in ("mass", "ke", "pe", "px", "py", "pz", "lx", "ly", "lz", "cx", "cy", "cz")]]
    for _ in range(10):
        R.kick(0.5 * dt)
        R.drift(dt)
        R.compute_forces(1.0, eps, 0.6, 0, True)
        R.kick(0.5 * dt)
        d = R.diagnostics()
        series.append([d[k] for k in ("mass", "ke", "pe", "px", "py", "pz", "lx", "ly", "lz", "cx", "cy", "cz")])
    fin = R.download()
    np.savez_compressed(os.path.join(HERE, "c0_direct.npz"), sample=sample, acc=acc[sample], pe=d0["pe"],
                        exact_pe=exact, series=np.array(series), final_ids=fin.id,
                        final_pos=np.stack([fin.x, fin.y, fin.z, fin.vx, fin.vy, fin.vz], 1)[np.argsort(fin.id)][sample])

    # 3. Tree, N = 2^20 Plummer (a = 1, cut 10), theta 0.6, eps 0.01.
    n1 = 1 << 20
    b = ref_plummer(n1, 1, 1.0, 10.0)
    r = tree_record(R, b, 0.01, 0.6, False)
    rng = np.random.default_rng(4096)
    s1 = np.sort(rng.choice(n1, 4096, replace=False)).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "c1_tree.npz"), sample=s1, acc=r["acc"][s1])
    out["c1_tree"] = dict(n=n1, eps=0.01, theta=0.6, interactions_listed=r["interactions"],
                          ids_digest=digest(r["ids"]), keys_digest=digest(r["keys"]),
                          canon_digest=digest(r["canon"]), nodes=int(len(r["canon"])),
                          leaves=int(r["canon"]["leaf"].sum()), costs_sum=float(r["costs"].astype(np.float64).sum()),
                          costs_digest=digest(r["costs"]))

    # 4. Small and degenerate sets, full reference outputs.
    small = {}
    cases = [(f"cube{n}", cube_bodies(n, 100 + n)) for n in (1, 2, 3, 31, 32, 33, 64, 255, 257, 1000, 1023)]
    cases.append(("coincident300", coincident_set()))
    for name, bb in cases:
        eps = 0.01 if name.startswith("plummer") else 0.05
        rec = {}
        for theta in ((0.0, 0.6) if len(bb) > 1 else (0.6,)):
            rr = tree_record(R, bb, eps, theta, True)
            d = R.diagnostics()
            key = f"t{theta}"
            rec[key] = rr
            rec[key]["pe"] = d["pe"]
        ad = R.accel_by_id(bb, 1.0, eps, 0.6, 0, True)
        R.upload(bb)
        R.compute_forces(1.0, eps, 0.6, 0, True)
        pe_direct = R.diagnostics()["pe"]
        small[name] = (bb, eps, rec, ad, pe_direct)
    arrays = {}
    meta = {}
    for name, (bb, eps, rec, ad, pe_direct) in small.items():
        arrays[f"{name}__bodies"] = np.stack([bb.x, bb.y, bb.z, bb.vx, bb.vy, bb.vz, bb.m], 1)
        arrays[f"{name}__direct_acc"] = ad
        m = dict(n=len(bb), eps=eps, pe_direct=pe_direct)
        for key, rr in rec.items():
            arrays[f"{name}__{key}__acc"] = rr["acc"]
            arrays[f"{name}__{key}__ids"] = rr["ids"]
            arrays[f"{name}__{key}__keys"] = rr["keys"]
            arrays[f"{name}__{key}__canon"] = rr["canon"]
            if rr["costs"] is not None:
                arrays[f"{name}__{key}__costs"] = rr["costs"]
            m[key] = dict(interactions_listed=rr["interactions"], pe=rr["pe"])
        meta[name] = m
    np.savez_compressed(os.path.join(HERE, "small_sets.npz"), **arrays)
    out["small_sets"] = meta
    with open(os.path.join(HERE, "reference.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote fixtures to", HERE)


LET_CASE = (4096, 5, 1.0, 10.0)  # Plummer n, seed, a, cut
LET_FAR_BOX = ([24.8, -0.2, -0.2], [25.2, 0.2, 0.2])


def zero_v(b: O.Bodies) -> O.Bodies:
    b = b.copy()
    b.vx[:] = 0
    b.vy[:] = 0
    b.vz[:] = 0
    return b


def let_arrays(t: O.Let, tag: str) -> dict:
    return {f"{tag}__nodes": t.nodes, f"{tag}__bodies": np.stack([t.bx, t.by, t.bz, t.bm], 1)}


def write_let():
    """LET fixtures (src/backend.hpp:114-136): a Plummer sphere split at x = 0 into two ranks; rank 1
    exports its LET for rank 0's box (tree at theta 0.6, and direct); rank 0 computes its forces with
    it. Outputs per method: the LET, rank 0's accelerations, costs (rows in rank 0's upload order),
    pe and interactions. Also the whole sphere's LETs for a far box and for a point box."""
    b = zero_v(ref_plummer(*LET_CASE))
    h0, h1 = b.subset(b.x < 0), b.subset(b.x >= 0)
    arrays = {}
    for method, tag in ((1, "tree"), (0, "direct")):
        r0, r1 = O.RefCpu(), O.RefCpu()
        r0.upload(h0)
        r1.upload(h1)
        r0.build_tree(1.0, 0.01, 0.6, method)
        r1.build_tree(1.0, 0.01, 0.6, method)
        box0 = r0.local_bounds()
        let = r1.export_let(box0, 1.0, 0.01, 0.6, method)
        r0.compute_forces_with(1.0, 0.01, 0.6, method, True, [let])
        pe = r0.diagnostics()["pe"]
        inter = r0.times()["interactions"]
        costs = r0.body_costs()
        r0.kick(1.0)
        o = r0.download()
        row = np.empty(int(h0.id.max()) + 1, np.int64)
        row[h0.id] = np.arange(len(h0))
        acc = np.zeros((len(h0), 3), np.float32)
        cst = np.zeros(len(h0), np.float32)
        acc[row[o.id]] = np.stack([o.vx, o.vy, o.vz], 1)
        cst[row[o.id]] = costs
        arrays.update(let_arrays(let, tag))
        arrays[f"{tag}__box0"] = np.array(box0, np.float32)
        arrays[f"{tag}__acc"] = acc
        arrays[f"{tag}__costs"] = cst
        arrays[f"{tag}__pe"] = np.float64(pe)
        arrays[f"{tag}__interactions"] = np.uint64(inter)
    R = O.RefCpu()
    R.upload(b)
    R.build_tree(1.0, 0.01, 0.6, 1)
    arrays.update(let_arrays(R.export_let(LET_FAR_BOX, 1.0, 0.01, 0.6, 1), "far"))
    p = [float(h0.x[0]), float(h0.y[0]), float(h0.z[0])]
    arrays.update(let_arrays(R.export_let((p, p), 1.0, 0.01, 0.6, 1), "point"))
    np.savez_compressed(os.path.join(HERE, "let.npz"), **arrays)
    print("wrote", os.path.join(HERE, "let.npz"))


if __name__ == "__main__":
    if sys.argv[1:] == ["--let"]:
        write_let()
    else:
        main()
        write_let()
