"""CPU tests: the oracle restatement pinned against the reference's golden fixtures, against the
reference build itself where oracle/_ref exists, and against closed forms and invariants."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import oracle as O
from helpers import canon_from_oracle, canon_from_ref_export, digest, morton_keys_np, rel_rms

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLD, "reference.json")))
SMALL = np.load(os.path.join(GOLD, "small_sets.npz"))


def bodies_from(arr) -> O.Bodies:
    n = len(arr)
    b = O.Bodies.empty(n)
    for k, a in enumerate((b.x, b.y, b.z, b.vx, b.vy, b.vz, b.m)):
        a[:] = arr[:, k]
    b.id[:] = np.arange(n, dtype=np.uint32)
    return b


# ---------------------------------------------------------------- scenario generator

@pytest.mark.parametrize("case", REF["plummer"], ids=lambda c: f"n{c['n']}_s{c['seed']}_a{c['a']:.3f}")
def test_plummer_generator_bit_exact(case):
    b = O.plummer(case["n"], case["seed"], case["a"], case["cut"])
    assert digest(*b.arrays()) == case["digest"]


def test_plummer_closed_form_moments():
    # Plummer a = 1: the half-mass radius is 1/sqrt(2^(2/3) - 1) = 1.3048 (cut at 10 removes 1.5%
    # of the mass, shifting it slightly inward); virial ratio 2K/|W| ~ 1.
    b = O.plummer(65536, 3, 1.0, 10.0)
    r = np.sqrt(b.x.astype(np.float64) ** 2 + b.y ** 2 + b.z ** 2)
    assert abs(np.median(r) - 1.3048) < 0.03
    assert abs(float(b.m.astype(np.float64).sum()) - 1.0) < 1e-6
    assert r.max() <= 10.0 + 1e-2  # cut at 10 before the centre-of-mass shift


# ---------------------------------------------------------------- keys and sort

def test_keys_known_answers():
    # The oracle's keys for bodies at exactly known quantised coordinates: origin = the float minimum
    # 0, side = 1 * (1 + 1e-6), so a coordinate v quantises to floor(v * 2^21 / 1.000001): 1 -> 2097149
    # (0x1ffffd) and 0.5 -> 1048574 (0xffffe). Interleaving x << 2 | y << 1 | z of the 21-bit
    # dilations by hand: (1, 1, 1) -> 7 * dilate(0x1ffffd) = 0x7fffffffffffffc7 (every bit but the
    # three of level bit 1); (0.5, 0.5, 0.5) -> 0x0ffffffffffffff8; (1, 0, 0.5) ->
    # dilate(0x1ffffd) << 2 | dilate(0xffffe) = 0x4b6db6db6db6db4c; the origin -> 0. The sorted order
    # is then bodies 0, 2, 3, 1.
    b = O.Bodies.empty(4)
    for k, (px, py, pz) in enumerate(((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (0.5, 0.5, 0.5), (1.0, 0.0, 0.5))):
        b.x[k], b.y[k], b.z[k] = px, py, pz
    b.m[:] = 0.25
    b.id[:] = np.arange(4, dtype=np.uint32)
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, 0.01, 0.6, 1)
    keys, perm = o.keys()
    assert [int(k) for k in keys] == [0, 0x0FFFFFFFFFFFFFF8, 0x4B6DB6DB6DB6DB4C, 0x7FFFFFFFFFFFFFC7]
    assert list(perm) == [0, 2, 3, 1]
    assert list(o.download().id) == [0, 2, 3, 1]


def test_keys_match_closed_form_and_stable_sort():
    b = O.plummer(50000, 11)
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, 0.01, 0.6, 1)
    keys, perm = o.keys()
    s = o.download()
    # keys of the sorted positions (same cube: the bounding box is order-independent)
    assert np.array_equal(keys, morton_keys_np(s.x, s.y, s.z))
    assert np.all(np.diff(keys.astype(np.uint64)) >= 0) or np.all(keys[1:] >= keys[:-1])
    # permutation = the stable argsort of the unsorted keys
    k0 = morton_keys_np(b.x, b.y, b.z)
    assert np.array_equal(perm, np.argsort(k0, kind="stable").astype(np.uint32))
    assert np.array_equal(s.id, b.id[perm])


def test_c1_keys_permutation_topology_match_reference():
    c = REF["c1_tree"]
    b = O.plummer(c["n"], 1, 1.0, 10.0)
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, c["eps"], c["theta"], 1)
    keys, _ = o.keys()
    s = o.download()
    assert digest(s.id) == c["ids_digest"]
    assert digest(keys) == c["keys_digest"]
    canon = canon_from_oracle(o.nodes())
    assert len(canon) == c["nodes"]
    assert int(canon["leaf"].sum()) == c["leaves"]
    assert digest(canon) == c["canon_digest"]
    cnt = o.counts()
    assert cnt["listed"] == c["interactions_listed"]
    costs = o.body_costs()
    assert digest(costs) == c["costs_digest"]
    g = np.load(os.path.join(GOLD, "c1_tree.npz"))
    ao = np.zeros((len(s), 3))
    ax, ay, az, _ = o.accel()
    ao[s.id, 0], ao[s.id, 1], ao[s.id, 2] = ax, ay, az
    assert rel_rms(ao[g["sample"]], g["acc"]) < 2e-6


# ---------------------------------------------------------------- octree invariants and moments

def test_octree_invariants_and_moments_closed_form():
    b = O.plummer(30000, 5)
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, 0.01, 0.6, 1)
    nd = o.nodes()
    s = o.download()
    cnt = nd["hi"] - nd["lo"]
    leaf = nd["nchild"] == 0
    assert np.all(cnt[leaf] <= 16) and np.all(cnt[~leaf] > 16)
    # children partition their parent's range in order
    for k in np.flatnonzero(~leaf)[:2000]:
        f, c = nd["first"][k], nd["nchild"][k]
        assert nd["lo"][f] == nd["lo"][k] and nd["hi"][f + c - 1] == nd["hi"][k]
        assert np.all(nd["hi"][f:f + c - 1] == nd["lo"][f + 1:f + c])
    # root mass is the total mass; quadrupoles are traceless
    assert abs(nd["m"][0] - float(s.m.astype(np.float64).sum())) < 1e-6
    q = nd["q"].astype(np.float64)
    scale = np.abs(q).max(axis=1) + 1e-30
    assert np.all(np.abs(q[:, 0] + q[:, 3] + q[:, 5]) <= 1e-5 * scale + 1e-12)
    # moments by brute force from the bodies of a node (com and Q about it, in double)
    rng = np.random.default_rng(0)
    for k in rng.choice(len(cnt), 200, replace=False):
        sl = slice(nd["lo"][k], nd["hi"][k])
        m = s.m[sl].astype(np.float64)
        p = np.stack([s.x[sl], s.y[sl], s.z[sl]], 1).astype(np.float64)
        M = m.sum()
        com = (m[:, None] * p).sum(0) / M
        d = p - com
        r2 = (d ** 2).sum(1)
        Q = 3 * np.einsum("i,ij,ik->jk", m, d, d) - np.eye(3) * (m * r2).sum()
        qv = np.array([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]])
        assert np.allclose(nd["com"][k], com, rtol=0, atol=2e-6 * (1 + np.abs(com).max()))
        assert abs(nd["m"][k] - M) <= 1e-6 * M
        assert np.allclose(nd["q"][k], qv, rtol=1e-4, atol=1e-4 * np.abs(qv).max() + 1e-12)
        # rcrit = 2 half / theta + |com - centre|
        g = nd["geo"][k].astype(np.float64)
        r = 2 * g[3] / 0.6 + np.linalg.norm(nd["com"][k].astype(np.float64) - g[:3])
        assert abs(nd["rcrit2"][k] - r * r) <= 1e-5 * r * r


def test_two_body_node_moments():
    b = O.Bodies.empty(2)
    b.x[:] = [0.0, 1.0]
    b.m[:] = [1.0, 3.0]
    b.id[:] = [0, 1]
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, 0.0, 0.5, 1)
    nd = o.nodes()
    assert len(nd["m"]) == 1  # two bodies: the root is a leaf
    assert nd["m"][0] == 4.0
    assert abs(nd["com"][0][0] - 0.75) < 1e-7
    # Q_xx = sum m (3 dx^2 - dx^2) = 2 * (1 * 0.75^2 + 3 * 0.25^2) = 1.5
    assert abs(nd["q"][0][0] - 1.5) < 1e-6 and abs(nd["q"][0][3] + 0.75) < 1e-6


# ---------------------------------------------------------------- forces

def test_direct_f64_two_body_closed_form():
    b = O.Bodies.empty(2)
    b.x[:] = [0.0, 1.0]
    b.m[:] = [1.0, 1.0]
    for eps in (0.0, 0.1):
        a, phi = O.direct_f64(b, 1.0, eps)
        assert abs(a[0, 0] - 1 / (1 + eps * eps) ** 1.5) < 1e-15
        assert abs(a[1, 0] + 1 / (1 + eps * eps) ** 1.5) < 1e-15
        assert abs(phi[0] + 1 / math.sqrt(1 + eps * eps)) < 1e-15


def test_oracle_direct_vs_fp64_and_reference_sample():
    g = np.load(os.path.join(GOLD, "c0_direct.npz"))
    b = O.plummer(16384, 1, 1.0, 10.0)
    o = O.Oracle()
    a = o.accel_by_id(b, 1.0, 0.01, 0.6, 0)
    a64, _ = O.direct_f64(b, 1.0, 0.01, g["sample"])
    assert rel_rms(a[g["sample"]], a64) < 1e-5  # the reference test's bound (tests/test_cpu_backend.cpp:126)
    assert rel_rms(a[g["sample"]], g["acc"]) < 5e-6
    assert o.counts()["interactions"] == 16384 * 16383


def test_tree_theta0_is_direct_and_counts():
    b = O.plummer(3000, 9)
    o = O.Oracle()
    ad = o.accel_by_id(b, 1.0, 0.05, 0.6, 0)
    at = o.accel_by_id(b, 1.0, 0.05, 0.0, 1)
    c = o.counts()
    assert rel_rms(at, ad) < 5e-6
    assert c["interactions"] == 3000 * 2999 and c["listed"] == 3000 * 3000


def test_tree_error_bounds_vs_fp64():
    # reference test: tree vs direct <= 1e-3 at theta 0.6 and <= 3e-4 at 0.4 (tests/test_cpu_backend.cpp:129-139)
    b = O.plummer(10000, 2, 3 * math.pi / 16, 20.0)
    o = O.Oracle()
    a64, _ = O.direct_f64(b, 1.0, 0.01)
    assert rel_rms(o.accel_by_id(b, 1.0, 0.01, 0.6, 1), a64) <= 1e-3
    assert rel_rms(o.accel_by_id(b, 1.0, 0.01, 0.4, 1), a64) <= 3e-4


@pytest.mark.parametrize("name", sorted(REF["small_sets"]))
def test_small_sets_match_reference(name):
    meta = REF["small_sets"][name]
    b = bodies_from(SMALL[f"{name}__bodies"])
    eps = meta["eps"]
    o = O.Oracle()
    ad = o.accel_by_id(b, 1.0, eps, 0.6, 0, True)
    assert rel_rms(ad, SMALL[f"{name}__direct_acc"]) < 5e-6
    for key in ("t0.0", "t0.6"):
        if key not in meta:
            continue
        theta = float(key[1:])
        a = o.accel_by_id(b, 1.0, eps, theta, 1, True)
        s = o.download()
        assert np.array_equal(s.id, SMALL[f"{name}__{key}__ids"])
        keys, _ = o.keys()
        if "coincident" not in name:  # restarted cells carry their own keys
            assert np.array_equal(keys, SMALL[f"{name}__{key}__keys"])
        assert np.array_equal(canon_from_oracle(o.nodes()), SMALL[f"{name}__{key}__canon"])
        assert o.counts()["listed"] == meta[key]["interactions_listed"]
        assert rel_rms(a, SMALL[f"{name}__{key}__acc"]) < 5e-6


# ---------------------------------------------------------------- integration and diagnostics

def test_kick_drift_diagnostics_closed_form():
    b = O.plummer(3000, 3)
    rng = np.random.default_rng(4)
    b.m *= rng.uniform(0.5, 1.5, len(b)).astype(np.float32)
    b.x += np.float32(0.3)
    b.vy += np.float32(0.2)
    o = O.Oracle()
    o.upload(b)
    d = o.diagnostics()
    m, x, y, z = (v.astype(np.float64) for v in (b.m, b.x, b.y, b.z))
    vx, vy, vz = (v.astype(np.float64) for v in (b.vx, b.vy, b.vz))
    assert abs(d["mass"] - m.sum()) < 1e-12
    assert abs(d["ke"] - 0.5 * (m * (vx * vx + vy * vy + vz * vz)).sum()) < 1e-12
    assert abs(d["lz"] - (m * (x * vy - y * vx)).sum()) < 1e-12
    assert abs(d["cx"] - (m * x).sum() / m.sum()) < 1e-12
    # kick: v += a dt as one fused multiply-add per component
    o.compute_forces(1.0, 0.01, 0.6, 0)
    ax, ay, az, _ = o.accel()
    o.kick(0.01)
    s = o.download()
    exp = (ax.astype(np.float64) * np.float64(np.float32(0.01)) + b.vx.astype(np.float64)).astype(np.float32)
    assert np.array_equal(s.vx, exp)


def test_c0_kdk_series_matches_reference():
    g = np.load(os.path.join(GOLD, "c0_direct.npz"))
    b = O.plummer(16384, 1, 1.0, 10.0)
    o = O.Oracle()
    o.upload(b)
    o.compute_forces(1.0, 0.01, 0.6, 0, True)
    rows = [o.diagnostics()]
    assert abs(o.exact_pe(1.0, 0.01) / float(g["exact_pe"]) - 1) < 1e-12
    for _ in range(10):
        o.kick(5e-4)
        o.drift(1e-3)
        o.compute_forces(1.0, 0.01, 0.6, 0, True)
        o.kick(5e-4)
        rows.append(o.diagnostics())
    ref = g["series"]
    e = np.array([r["ke"] + r["pe"] for r in rows])
    e_ref = ref[:, 1] + ref[:, 2]
    assert np.max(np.abs(e / e_ref - 1)) < 1e-6
    assert np.max(np.abs(np.array([r["ke"] for r in rows]) / ref[:, 1] - 1)) < 1e-5


# ---------------------------------------------------------------- new generator (no reference counterpart)

def test_disk_sech2_closed_form_moments():
    # Pinned by the profile's closed forms: E|z| = z0 ln 2 for sech^2(z/z0); radius of a truncated
    # exponential disk: E[R] = Rd * g(3, x) / g(2, x) with g the lower incomplete gamma, x = rmax/Rd.
    z0, rd, rmax = 0.1, 1.0, 10.0
    b = O.disk_sech2(200000, seed=5, core=10.0, z0=z0, rd=rd, rmax=rmax, com_frame=False)
    z = b.z[1:].astype(np.float64)
    R = np.hypot(b.x[1:].astype(np.float64), b.y[1:].astype(np.float64))
    assert abs(np.abs(z).mean() - z0 * math.log(2)) < 4 * z0 * 0.9 / math.sqrt(len(z))
    x = rmax / rd
    g2 = 1 - math.exp(-x) * (1 + x)
    g3 = 2 * (1 - math.exp(-x) * (1 + x + x * x / 2))
    assert abs(R.mean() - rd * g3 / g2) < 4 * 1.5 / math.sqrt(len(R))
    assert b.m[0] == 10.0 and abs(float(b.m[1:].astype(np.float64).sum()) - 1.0) < 1e-4
    # circular speed from the enclosed mass, dispersion 0.1 vc per component
    vphi = (b.x[1:] * b.vy[1:] - b.y[1:] * b.vx[1:]).astype(np.float64) / R
    order = np.argsort(R)
    enc = 10.0 + np.arange(1, len(R) + 1) / len(R)
    vc = np.sqrt(enc / R[order])
    assert abs(np.mean(vphi[order] / vc) - 1) < 0.01


# ---------------------------------------------------------------- locally essential trees

LETZ = np.load(os.path.join(GOLD, "let.npz"))


def zero_v(b: O.Bodies) -> O.Bodies:
    b = b.copy()
    b.vx[:] = 0
    b.vy[:] = 0
    b.vz[:] = 0
    return b


def let_halves():
    """The LET fixture's inputs: a 4096-body Plummer sphere (seed 5) split at x = 0."""
    b = zero_v(O.plummer(4096, 5))
    return b, b.subset(b.x < 0), b.subset(b.x >= 0)


def let_body_table(t) -> np.ndarray:
    return np.stack([t.bx, t.by, t.bz, t.bm], 1)


def oracle_accel(o: O.Oracle, rows_of_id, n):
    """Accelerations (and costs) after a force call on bodies uploaded with v = 0, by upload row."""
    costs = o.body_costs()
    o.kick(1.0)
    d = o.download()
    a = np.zeros((n, 3), np.float32)
    c = np.zeros(n, np.float32)
    a[rows_of_id[d.id]] = np.stack([d.vx, d.vy, d.vz], 1)
    c[rows_of_id[d.id]] = costs
    return a, c


@pytest.mark.parametrize("method,tag", [(1, "tree"), (0, "direct")])
def test_let_export_and_remote_walk_match_reference_fixture(method, tag):
    # Two ranks (src/backend.hpp:114-136): rank 1's LET for rank 0's box is the reference's record for
    # record, and rank 0's forces with it have the reference's lists (costs, counts) and values.
    _, h0, h1 = let_halves()
    o0, o1 = O.Oracle(), O.Oracle()
    o0.upload(h0)
    o1.upload(h1)
    o0.build_tree(0.6, method)
    o1.build_tree(0.6, method)
    box0 = o0.local_bounds()
    assert np.array_equal(np.array(box0, np.float32), LETZ[f"{tag}__box0"])
    let = o1.export_let(box0, method)
    assert let.nodes.tobytes() == LETZ[f"{tag}__nodes"].tobytes()
    assert np.array_equal(let_body_table(let), LETZ[f"{tag}__bodies"])
    o0.compute_forces_with(1.0, 0.01, method, True, [let])
    pe = o0.diagnostics()["pe"]
    c = o0.counts()
    rows = np.zeros(4096, np.int64)
    rows[h0.id] = np.arange(len(h0))
    a, costs = oracle_accel(o0, rows, len(h0))
    assert np.array_equal(costs, LETZ[f"{tag}__costs"])
    assert c["listed"] == int(LETZ[f"{tag}__interactions"])
    assert rel_rms(a, LETZ[f"{tag}__acc"]) < 5e-6
    assert abs(pe / float(LETZ[f"{tag}__pe"]) - 1) < 1e-6


def test_let_far_and_point_boxes_match_reference_fixture():
    b, h0, _ = let_halves()
    o = O.Oracle()
    o.upload(b)
    o.build_tree(0.6, 1)
    far = o.export_let(([24.8, -0.2, -0.2], [25.2, 0.2, 0.2]), 1)
    assert far.nodes.tobytes() == LETZ["far__nodes"].tobytes() and len(far.bx) == 0
    p = [float(h0.x[0]), float(h0.y[0]), float(h0.z[0])]
    pt = o.export_let((p, p), 1)
    assert pt.nodes.tobytes() == LETZ["point__nodes"].tobytes()
    assert np.array_equal(let_body_table(pt), LETZ["point__bodies"])


def test_let_box_holding_everything_exports_the_whole_tree():
    # Every node's com lies in the box, so no node is terminal: the LET is the whole tree (every
    # node has positive mass here) and every body, leaf by leaf in breadth-first order.
    b = zero_v(O.plummer(3000, 8))
    o = O.Oracle()
    o.upload(b)
    o.build_tree(0.6, 1)
    lo, hi = o.local_bounds()
    let = o.export_let((lo, hi), 1)
    nodes = o.nodes()
    assert len(let.nodes) == len(nodes["m"])
    assert len(let.bx) == len(b)
    leaf = let.nodes["body_count"] > 0
    assert np.array_equal(leaf, let.nodes["first_child"] < 0)
    assert let.nodes["body_count"].sum() == len(b)
    assert np.array_equal(np.sort(let.bm), np.sort(b.m)) and np.isclose(let.nodes["mass"][0], 1.0, rtol=1e-6)
    # children follow their parent, siblings contiguous and in order
    for k in np.flatnonzero(~leaf):
        f, c = let.nodes["first_child"][k], let.nodes["child_count"][k]
        assert f > k and c >= 1 and np.isclose(let.nodes["mass"][f:f + c].sum(), let.nodes["mass"][k], rtol=1e-5)


def test_let_theta0_two_ranks_equal_one_direct_sum():
    # theta 0 opens every remote node: the two ranks' forces are the direct sum over all bodies, and
    # the pair counts add up to n (n - 1).
    b = zero_v(O.plummer(2000, 9))
    h0, h1 = b.subset(b.x < 0.1), b.subset(b.x >= 0.1)
    ad = O.Oracle().accel_by_id(b, 1.0, 0.01, 0.6, 0)
    a = np.zeros((len(b), 3))
    inter = 0
    for method in (1, 0):
        o = [O.Oracle(), O.Oracle()]
        for k, h in enumerate((h0, h1)):
            o[k].upload(h)
            o[k].build_tree(0.0, method)
        lets = [o[1].export_let(o[0].local_bounds(), method), o[0].export_let(o[1].local_bounds(), method)]
        inter = 0
        for k in range(2):
            o[k].compute_forces_with(1.0, 0.01, method, False, [lets[k]])
            inter += o[k].counts()["interactions"]
            o[k].kick(1.0)
            d = o[k].download()
            a[d.id] = np.stack([d.vx, d.vy, d.vz], 1)
        assert inter == len(b) * (len(b) - 1)
        assert rel_rms(a, ad) < 1e-6


def test_let_empty_and_zero_mass():
    # An empty LET changes nothing; a root of zero mass exports nothing; zero-mass children are left out.
    b = zero_v(O.plummer(3000, 10))
    o = O.Oracle()
    a0 = o.accel_by_id(b, 1.0, 0.01, 0.6, 1, True)
    o.upload(b)
    o.build_tree(0.6, 1)
    empty = O.Let.sized(0, 0)
    o.compute_forces_with(1.0, 0.01, 1, True, [empty, empty])
    o.kick(1.0)
    d = o.download()
    a1 = np.zeros_like(a0)
    a1[d.id] = np.stack([d.vx, d.vy, d.vz], 1)
    assert np.array_equal(a0, a1)
    z = b.copy()
    z.m[:] = 0
    o.upload(z)
    o.build_tree(0.6, 1)
    assert len(o.export_let(([-1, -1, -1], [1, 1, 1]), 1).nodes) == 0
    h = b.copy()
    h.m[h.x > 0] = 0  # every cell on the x > 0 side is massless
    o.upload(h)
    o.build_tree(0.6, 1)
    let = o.export_let(([-100, -100, -100], [100, 100, 100]), 1)
    assert (let.nodes["mass"] > 0).all()
    leaves = np.flatnonzero(let.nodes["body_count"] > 0)
    for k in leaves:  # every exported leaf holds mass; massless leaves stay home
        f, c = let.nodes["body_first"][k], let.nodes["body_count"][k]
        assert let.bm[f:f + c].sum() > 0
    assert len(let.bx) < len(b) and (let.bm > 0).sum() == (h.m > 0).sum()


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("method", [1, 0])
def test_oracle_let_matches_reference_build_live(method):
    # Random halves (overlapping boxes) of a 20k sphere, both directions, with the potential.
    b = zero_v(O.plummer(20000, 13))
    pick = np.random.default_rng(99).integers(0, 2, len(b)).astype(bool)
    halves = (b.subset(pick), b.subset(~pick))
    o = [O.Oracle(), O.Oracle()]
    r = [O.RefCpu(), O.RefCpu()]
    for k in range(2):
        o[k].upload(halves[k])
        r[k].upload(halves[k])
        o[k].build_tree(0.6, method)
        r[k].build_tree(1.0, 0.01, 0.6, method)
    for k in range(2):
        box = r[k].local_bounds()
        assert o[k].local_bounds() == box
        lo_ = o[1 - k].export_let(box, method)
        lr = r[1 - k].export_let(box, 1.0, 0.01, 0.6, method)
        assert lo_.nodes.tobytes() == lr.nodes.tobytes()
        assert np.array_equal(let_body_table(lo_), let_body_table(lr))
        o[k].compute_forces_with(1.0, 0.01, method, True, [lo_])
        r[k].compute_forces_with(1.0, 0.01, 0.6, method, True, [lr])
        assert np.array_equal(o[k].body_costs(), r[k].body_costs())
        assert o[k].counts()["listed"] == r[k].times()["interactions"]
        assert abs(o[k].diagnostics()["pe"] / r[k].diagnostics()["pe"] - 1) < 1e-6
        o[k].kick(1.0)
        r[k].kick(1.0)
        do, dr = o[k].download(), r[k].download()
        assert np.array_equal(do.id, dr.id)
        assert rel_rms(np.stack([do.vx, do.vy, do.vz], 1), np.stack([dr.vx, dr.vy, dr.vz], 1)) < 5e-6


# ---------------------------------------------------------------- live reference (this container only)

@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("theta", [0.6, 0.3, 0.0])
def test_oracle_matches_reference_build_live(theta):
    b = O.plummer(20000, 4)
    o, R = O.Oracle(), O.RefCpu()
    ao = o.accel_by_id(b, 1.0, 0.01, theta, 1)
    ar = R.accel_by_id(b, 1.0, 0.01, theta, 1)
    assert np.array_equal(o.download().id, R.download().id)
    assert np.array_equal(o.body_costs(), R.body_costs())
    assert o.counts()["listed"] == R.times()["interactions"]
    assert rel_rms(ao, ar) < 5e-6
    t = R.export_tree(1.0, 0.01, theta)
    o.compute_forces(1.0, 0.01, theta, 1)
    assert np.array_equal(canon_from_oracle(o.nodes()), canon_from_ref_export(t))


# ---------------------------------------------------------------- FP64 adjoint (no reference counterpart)

def test_oracle_adjoint_matches_central_differences():
    # Pins orc_sim_backward against the mathematics: central differences of orc_sim_forward for
    # positions, velocities and masses, loss L = sum |x_T - x*|^2 + sum v_T^2 / 3. The horizon is one
    # crossing time (200 steps of 5e-3), so the force Jacobian carries a large share of each gradient
    # (the part of dL/dx0 beyond the identity term 2 (x_T - x*) is ~2/3 of its rms): a wrong sign,
    # a dropped term or a transposed operand in the adjoint fails by orders of magnitude.
    b = O.plummer(64, 5)
    x0 = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    v0 = np.stack([b.vx, b.vy, b.vz], 1).astype(np.float64)
    m = b.m.astype(np.float64) * (1 + 0.01 * np.random.default_rng(5).standard_normal(len(b)))
    G, eps, dt, steps = 1.0, 0.05, 5e-3, 200
    tgt = x0 + 0.1 * np.random.default_rng(11).standard_normal(x0.shape)

    def loss(x0_, v0_, m_):
        xT, vT = O.sim_forward(x0_, v0_, m_, G, eps, dt, steps)
        return float(((xT - tgt) ** 2).sum() + (vT ** 2).sum() / 3)

    xT, vT = O.sim_forward(x0, v0, m, G, eps, dt, steps)
    gx, gv, gm = O.sim_backward(2 * (xT - tgt), 2 * vT / 3)
    force_part = np.sqrt(((gx - 2 * (xT - tgt)) ** 2).mean())
    assert force_part > 0.3 * np.sqrt((gx ** 2).mean())
    h = 1e-5
    rng = np.random.default_rng(3)
    for _ in range(8):
        i, c = int(rng.integers(len(m))), int(rng.integers(3))
        for arr, g in ((x0, gx), (v0, gv)):
            ap, am = arr.copy(), arr.copy()
            ap[i, c] += h
            am[i, c] -= h
            args_p = (ap, v0, m) if arr is x0 else (x0, ap, m)
            args_m = (am, v0, m) if arr is x0 else (x0, am, m)
            fd = (loss(*args_p) - loss(*args_m)) / (2 * h)
            floor = 0.1 * np.sqrt((g ** 2).mean())
            assert abs(fd - g[i, c]) <= 1e-6 * max(floor, abs(fd)), (fd, g[i, c])
        mp, mm = m.copy(), m.copy()
        mp[i] += h * m[i]
        mm[i] -= h * m[i]
        fd = (loss(x0, v0, mp) - loss(x0, v0, mm)) / (2 * h * m[i])
        assert abs(fd - gm[i]) <= 1e-6 * max(0.1 * np.sqrt((gm ** 2).mean()), abs(fd)), (fd, gm[i])
