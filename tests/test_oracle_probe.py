"""Pins of the oracle's probe sums, SmoothedParticles.sum(sys, func, x)
(src/core.jl:240-258) with the interpolation closures of
examples/cavity_flow.jl:166-172.  Each check is independent of the oracle's
own loop: a numpy brute force over all particles, the oracle's density sum
(pinned by lattice sums elsewhere), closed forms, and the reference's NaN
rule worked out from the key geometry.  CPU only."""
import math

import numpy as np
import pytest

import oracle as O

L = O.lib()


def cloud(n, dim, seed, h):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    lo = np.zeros(3)
    hi = np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0])
    o = O.OracleSystem(h, lo, hi)
    v = rng.normal(size=(n, 3))
    if dim == 2:
        v[:, 2] = 0.0
    typ = (rng.random(n) < 0.3).astype(float)
    o.load(x, v=v, type=typ)
    assert o.create_cell_list() == n
    return o, x, v, typ, rng


def W(dim, h, r):
    return (L.orc_wendland2 if dim == 2 else L.orc_wendland3)(h, r)


def brute(x, v, typ, m, h, dim, p, fluid_only):
    """Sum over every particle with RN(sqrt((d1*d1 + d2*d2) + d3*d3)) <= h."""
    d = p[None, :] - x
    r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    out = np.zeros(4)
    for q in np.nonzero(r <= h)[0]:
        w = (1.0 if (not fluid_only or typ[q] == 0.0) else 0.0) * m
        k = W(dim, h, float(r[q]))
        out[0] += w * k
        out[1:] += w * v[q] * k
    return out


@pytest.mark.parametrize("dim,n,h", [(2, 1500, 0.06), (3, 2500, 0.15)])
@pytest.mark.parametrize("fluid_only", [False, True])
def test_probe_sum_brute_force(dim, n, h, fluid_only):
    """Probes inside the box of a grid that does not alias: the reference
    visits every particle within h, so the sums equal an O(N) brute force up
    to summation order."""
    o, x, v, typ, rng = cloud(n, dim, 3 + n, h)
    P = rng.random((40, 3)) * 0.9 + 0.05
    if dim == 2:
        P[:, 2] = 0.0
    P[:5] = x[:5]                      # at particles (r = 0 term)
    P[5] = [2 * h, 3 * h, 0.0 if dim == 2 else 2 * h]  # on cell corners
    got = o.probe_sum(P, 0.7, fluid_only)
    for i, p in enumerate(P):
        ref = brute(x, v, typ, 0.7, h, dim, p, fluid_only)
        np.testing.assert_allclose(got[i], ref, rtol=1e-12, atol=1e-12 * abs(ref[0]))


@pytest.mark.parametrize("dim,n,h", [(2, 1200, 0.06), (3, 1500, 0.15)])
def test_probe_sum_visits_exactly_the_neighbours(dim, n, h):
    """With v_x = 1/W(r) for the probe's particles (0 elsewhere) and m = 1,
    the x sum counts the visited particles with W > 0; it must equal the
    brute-force count, so no particle near r = h is missed or added."""
    o, x, v, typ, rng = cloud(n, dim, 7 + n, h)
    for t in range(12):
        p = rng.random(3) * 0.9 + 0.05
        if dim == 2:
            p[2] = 0.0
        d = p[None, :] - x
        r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        vv = np.zeros((n, 3))
        keep = 0
        for q in np.nonzero(r <= h)[0]:
            k = W(dim, h, float(r[q]))
            if k > 0.0:
                vv[q, 0] = 1.0 / k
                keep += 1
        o.field("v")[:] = vv
        got = o.probe_sum(p, 1.0)[0]
        assert abs(got[1] - keep) < 1e-9 * max(keep, 1)
        assert keep > 0


@pytest.mark.parametrize("dim", [2, 3])
def test_probe_at_particle_equals_density_sum(dim):
    """At a particle's own position the probe visits that particle's
    neighbours and itself (r = 0): the density sum with the self term
    (find_rho! with self=true), up to the order of the self term."""
    h = 0.06 if dim == 2 else 0.15
    o, x, v, typ, _ = cloud(1000, dim, 21, h)
    m = 0.3
    o.reset_rho()
    o.density_sum(self_term=True, m=m)
    rho = o.field("rho").copy()
    got = o.probe_sum(x[:200], m)
    np.testing.assert_allclose(got[:, 0], rho[:200], rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_probe_sum_closed_forms(dim):
    """One particle at distance r: out = (m W(r), m v W(r)) exactly; a
    uniform velocity field interpolates to itself (Shepard ratio)."""
    h = 0.1
    hi = np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0])
    o = O.OracleSystem(h, np.zeros(3), hi)
    q = np.array([[0.5, 0.5, 0.5 if dim == 3 else 0.0]])
    o.load(q, v=np.array([[1.5, -2.0, 0.25 if dim == 3 else 0.0]]))
    o.create_cell_list()
    for r in (0.0, 0.03, 0.07, 0.0999, h, 0.1000001):
        p = q[0] + np.array([r, 0.0, 0.0])
        got = o.probe_sum(p, 0.8)[0]
        ra = math.sqrt((p[0] - q[0, 0]) * (p[0] - q[0, 0]))  # the distance as stored
        k = W(dim, h, ra) if ra <= h else 0.0
        assert got[0] == 0.8 * k
        assert got[1] == (0.8 * 1.5) * k and got[2] == (0.8 * -2.0) * k
    # uniform field
    o2, x, v, typ, rng = cloud(1500, dim, 5, h)
    v0 = np.array([0.3, -1.1, 0.7 if dim == 3 else 0.0])
    o2.field("v")[:] = v0
    P = rng.random((20, 3)) * 0.8 + 0.1
    if dim == 2:
        P[:, 2] = 0.0
    got = o2.probe_sum(P, 1.0)
    assert np.all(got[:, 0] > 0)
    np.testing.assert_allclose(got[:, 1:] / got[:, :1], np.broadcast_to(v0, (20, 3)),
                               rtol=1e-13, atol=1e-15)


def test_probe_sum_nonfinite_probes():
    """src/structs.jl:103-105: a NaN or infinite coordinate gives key -1.
    The reference then visits cells -1 + key_diff that pass the bounds test;
    a NaN distance is not rejected by r > h, so the sum is NaN exactly when
    one of those cells holds a particle.  An infinite coordinate gives
    r = Inf, rejected: the sum is 0."""
    h = 0.1
    o, x, v, typ, _ = cloud(800, 2, 9, h)
    _, _, kmax, diff = o.key_geometry()
    keys = np.array([o.find_key(xx) for xx in x])
    visited = [k for k in (-1 + np.array(diff)) if 1 <= k <= kmax]
    expect_nan = bool(np.isin(keys, visited).any())
    got = o.probe_sum(np.array([[math.nan, 0.5, 0.0], [math.inf, 0.5, 0.0],
                                [0.5, -math.inf, 0.0]]), 1.0)
    assert bool(np.isnan(got[0, 0])) == expect_nan
    assert expect_nan  # the cloud fills the grid, so the quirk is exercised
    assert np.all(got[1:] == 0.0)


def test_probe_sum_fluid_only_equals_fluid_subsystem():
    """phi = 0 for walls adds exact zeros, so the fluid-only sums are
    bit-identical to the unfiltered sums of a system holding the fluid
    particles only, in the same relative order."""
    h = 0.06
    o, x, v, typ, rng = cloud(1500, 2, 13, h)
    f = typ == 0.0
    o2 = O.OracleSystem(h, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o2.load(x[f], v=v[f])
    o2.create_cell_list()
    P = rng.random((30, 3))
    P[:, 2] = 0.0
    assert np.array_equal(o.probe_sum(P, 0.5, fluid_only=True), o2.probe_sum(P, 0.5))


# ---- SmoothedParticles.sum(sys, func, p), the per-particle form ------------

def lj_brute(x, typ, h, m, dr_wall, E_wall, eps):
    """All pairs (q == p included) with the FP64 predicate, math.fsum."""
    out = []
    for p_ in range(len(x)):
        terms = []
        for q_ in range(len(x)):
            d = x[p_] - x[q_]
            r = math.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
            if r > h:
                continue
            if typ[q_] == 1.0 and typ[p_] == 0.0 and r < dr_wall:
                s = dr_wall / (r + eps)
                terms.append(m * E_wall * (0.5 * s ** 2 - 0.25 * s ** 4 - 0.25))
        out.append(math.fsum(terms))
    return np.array(out)


def test_particle_sum_lj_brute_force_and_closed_form():
    """The wall potential of examples/collapse_symplectic.jl:144-162 summed
    per particle: against an all-pairs brute force (p == q included,
    math.fsum), and a single fluid-wall pair in closed form; walls and
    fluid-fluid pairs contribute nothing."""
    rng = np.random.default_rng(21)
    n = 900
    x = rng.random((n, 3)) * [1.0, 1.0, 0.0]
    typ = (rng.random(n) < 0.3).astype(np.float64)
    h, m, dr_wall, E_wall, eps = 0.06, 0.4, 0.03, 137.2, 1e-16
    o = O.OracleSystem(h, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o.load(x, type=typ)
    o.create_cell_list()
    got = o.particle_sum_lj(m, dr_wall, E_wall, eps)
    ref = lj_brute(x, typ, h, m, dr_wall, E_wall, eps)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert np.all(got[typ == 1.0] == 0.0) and np.count_nonzero(got) > 20
    # one fluid particle, one wall particle at r = 0.02
    x2 = np.array([[0.5, 0.5, 0.0], [0.52, 0.5, 0.0]])
    o2 = O.OracleSystem(h, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o2.load(x2, type=np.array([0.0, 1.0]))
    o2.create_cell_list()
    s = dr_wall / (0.02 + eps)
    v = o2.particle_sum_lj(m, dr_wall, E_wall, eps)
    assert v[0] == pytest.approx(m * E_wall * (0.5 * s * s - 0.25 * s ** 4 - 0.25), rel=1e-14)
    assert v[1] == 0.0
    # at r >= dr_wall nothing, and the potential is zero at r = dr_wall - eps
    o2.field("x")[1, 0] = 0.5 + dr_wall
    o2.create_cell_list()
    assert o2.particle_sum_lj(m, dr_wall, E_wall, eps)[0] == 0.0
