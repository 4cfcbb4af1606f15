"""CPU tests of the library's initial-condition generators (host code, no device needed): the
Plummer sphere against the reference generator's golden digests, the sech^2 disk against the
oracle's restatement and the closed forms of its profiles."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import lettuce_simulate_physics_amd as L
import oracle as O
from helpers import digest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = json.load(open(os.path.join(GOLD, "reference.json")))


@pytest.mark.parametrize("case", REF["plummer"], ids=lambda c: f"n{c['n']}_s{c['seed']}_a{c['a']:.3f}")
def test_plummer_bit_exact_with_the_reference_generator(case):
    b = L.plummer(case["n"], case["seed"], case["a"], case["cut"])
    assert digest(*b.arrays()) == case["digest"]


def test_plummer_f64_rounds_to_the_float_bodies():
    b = L.plummer(5000, 7)
    x, v, m = L.plummer_f64(5000, 7)
    assert np.array_equal(x[:, 0].astype(np.float32), b.x) and np.array_equal(v[:, 2].astype(np.float32), b.vz)
    assert np.array_equal(m.astype(np.float32), b.m)
    # centre-of-mass frame in double
    assert np.abs((x * m[:, None]).sum(0)).max() < 1e-12 and np.abs((v * m[:, None]).sum(0)).max() < 1e-12


@pytest.mark.parametrize("kw", [dict(), dict(incl=np.pi / 6, pos=(3.0, -2.0, 1.0), vel=(0.1, 0.2, -0.3),
                                             com_frame=False, spin=-1.0)])
def test_disk_matches_the_oracle_restatement(kw):
    a = L.disk_sech2(20000, seed=4, **kw)
    b = O.disk_sech2(20000, seed=4, **kw)
    for u, w in zip(a.arrays(), b.arrays()):
        assert np.array_equal(u, w)


def test_disk_closed_form_profiles():
    # sech^2(z/z0): E|z| = z0 ln 2; exponential surface density cut at rmax: mean radius
    # Rd g(3, x) / g(2, x) with g the lower incomplete gamma function, x = rmax / Rd.
    from scipy.special import gammainc, gamma
    z0, rd, rmax = 0.1, 1.0, 10.0
    x, v, m = L.disk_sech2_f64(200000, seed=5, z0=z0, rd=rd, rmax=rmax, com_frame=False)
    assert abs(np.mean(np.abs(x[1:, 2])) - z0 * np.log(2)) < 0.002
    R = np.hypot(x[1:, 0], x[1:, 1])
    xm = rmax / rd
    expect = rd * gammainc(3, xm) * gamma(3) / (gammainc(2, xm) * gamma(2))
    assert abs(R.mean() - expect) < 0.01
    assert m[0] == 10.0 and abs(m[1:].sum() - 1.0) < 1e-9
    # rotation: positive angular momentum about +z for spin = +1
    assert np.sum(m[1:] * (x[1:, 0] * v[1:, 1] - x[1:, 1] * v[1:, 0])) > 0


def test_generator_rejects_invalid_parameters():
    with pytest.raises(ValueError):
        L.plummer(10, 1, a=-1.0)
    with pytest.raises(ValueError):
        L.disk_sech2(0)


def test_collision_orbit_and_vjp_against_finite_differences():
    C = L.Collision(nstars=2000, seed=3, pericentre=4.0, incl1=0.5, incl2=1.0)
    x, v, m = C.generate()
    n1 = C.nstars + 1
    # cores: separation 20, relative energy zero (parabolic), relative angular momentum sqrt(2 G Mt q)
    r = x[n1] - x[0]
    u = v[n1] - v[0]
    Mt = 2 * (C.core + C.mass)
    assert abs(np.linalg.norm(r) - 20.0) < 1e-12
    assert abs(0.5 * u @ u - Mt / np.linalg.norm(r)) < 1e-12
    assert abs(np.linalg.norm(np.cross(r, u)) - np.sqrt(2 * Mt * 4.0)) < 1e-12
    assert r @ u < 0  # approaching
    # disk 2's spin axis is tilted by incl2 from +z about x
    d = x[n1 + 1:] - x[n1]
    w = v[n1 + 1:] - v[n1]
    Lz = (m[n1 + 1:, None] * np.cross(d, w)).sum(0)
    assert abs(np.arctan2(-Lz[1], Lz[2]) - 1.0) < 0.02
    # VJP against central differences of a random linear functional
    rng = np.random.default_rng(1)
    gx, gv = rng.standard_normal(x.shape), rng.standard_normal(v.shape)
    g = C.vjp(gx, gv)
    for k, name in enumerate(("pericentre", "incl1", "incl2")):
        h = 1e-6
        a, b = L.Collision(**{**C.__dict__, name: getattr(C, name) + h}), L.Collision(**{**C.__dict__, name: getattr(C, name) - h})
        xa, va, _ = a.generate()
        xb, vb, _ = b.generate()
        fd = ((gx * (xa - xb)).sum() + (gv * (va - vb)).sum()) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, g[k])
