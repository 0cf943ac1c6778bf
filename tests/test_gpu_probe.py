"""Probe sums on the GPU (sph_probe_sum, SmoothedParticles.sum at points,
src/core.jl:240-258) against the oracle.  Both walk the cells in key_diff
order and each cell in its stored order and evaluate the same FP64
expressions without contraction, so the comparison is bit-exact (NaN where
the reference gives NaN), in either precision mode of the handle."""
import math

import numpy as np
import pytest

import oracle as O
import smoothedparticles_hip as sp

pytestmark = pytest.mark.gpu

PHYS = dict(m=0.4, rho0=1000.0, c=50.0, alpha=0.02, hs=0.05, mu=0.0, nu=0.0,
            g=(0.0, -9.8, 0.0), dt=1e-4)


def make_both(x, h, lo, hi, precision, v, typ):
    g = sp.ParticleSystem(sp.Box(*lo, *hi), h, precision)
    g.add_particles(x, v=v, type=typ)
    o = O.OracleSystem(h, np.asarray(lo, float), np.asarray(hi, float))
    o.load(x, v, None, None, typ)
    return g, o


def probes(rng, x, h, dim, hi):
    P = rng.random((300, 3)) * hi * 1.1 - 0.05 * hi   # inside and just outside the box
    P[:40] = x[:40]                                  # at particles
    P[40:60] = np.round(P[40:60] / h) * h            # on cell faces and corners
    P[60] = [math.nan, 0.5, 0.0]
    P[61] = [math.inf, 0.5, 0.0]
    P[62] = [0.5, -math.inf, 0.0]
    P[63] = [0.2, math.nan, math.nan]
    P[64] = [5.0, 5.0, 5.0]                          # far outside
    if dim == 2:
        P[:60, 2] = 0.0
        P[65:250, 2] = 0.0                           # the rest probe off the plane
    return P


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim,n,h", [(2, 3000, 0.05), (3, 4000, 0.12), (2, 1500, 0.3)])
def test_probe_sums_bit_exact(precision, dim, n, h):
    rng = np.random.default_rng(31 + n)
    x = rng.random((n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    hi = np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0])
    v = rng.normal(size=(n, 3))
    typ = (rng.random(n) < 0.25).astype(float)
    g, o = make_both(x, h, np.zeros(3), hi, precision, v, typ)
    assert sp.create_cell_list(g) == o.create_cell_list() == n
    P = probes(rng, x, h, dim, np.where(hi > 0, hi, 1.0))
    for fluid_only in (False, True):
        got = sp.probe_sums(g, P, fluid_only=fluid_only, **PHYS)
        ref = o.probe_sum(P, PHYS["m"], fluid_only)
        assert np.array_equal(got, ref, equal_nan=True), (fluid_only, np.nanmax(abs(got - ref)))
    assert np.isnan(ref[60, 0]) or np.isnan(ref[63, 0]) or h > 0.2
    assert np.all(got[61:63] == 0.0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_probe_sums_after_steps(precision):
    """After device-resident steps and a rebuild, the probes read the moved
    particles through the new cell list; the oracle is loaded with the
    GPU's state, so the comparison stays bit-exact."""
    rng = np.random.default_rng(4)
    n, h = 2500, 0.05
    x = rng.random((n, 3)) * 0.8 + 0.1
    x[:, 2] = 0.0
    hi = np.array([1.0, 1.0, 0.0])
    g = sp.ParticleSystem(sp.Box(0, 0, 0, 1, 1, 0), h, precision)
    g.add_particles(x, v=np.zeros((n, 3)), type=(rng.random(n) < 0.2).astype(float))
    sp.step(g, 5, **PHYS)
    sp.create_cell_list(g)
    xs = g.get("x")
    o = O.OracleSystem(h, np.zeros(3), hi)
    o.load(xs, g.get("v"), None, None, g.get("type"))
    assert o.create_cell_list() == len(xs)
    P = probes(rng, xs, h, 2, np.array([1.0, 1.0, 1.0]))
    got = sp.probe_sums(g, P, fluid_only=True, **PHYS)
    ref = o.probe_sum(P, PHYS["m"], True)
    assert np.array_equal(got, ref, equal_nan=True)


def test_probe_sums_state_and_errors():
    """No current cell list: SPH_ESTATE; unknown flags: SPH_EINVAL; zero
    probes: nothing to do."""
    import ctypes
    from smoothedparticles_hip import _native as N
    rng = np.random.default_rng(2)
    x = rng.random((500, 3))
    x[:, 2] = 0.0
    g = sp.ParticleSystem(sp.Box(0, 0, 0, 1, 1, 0), 0.1, "fp64")
    g.add_particles(x)
    with pytest.raises(sp.SPHError):
        sp.probe_sums(g, [[0.5, 0.5, 0.0]], **PHYS)
    sp.create_cell_list(g)
    assert sp.probe_sums(g, np.zeros((0, 3)), **PHYS).shape == (0, 4)
    p = sp.make_params(**PHYS)
    buf = (ctypes.c_double * 4)()
    pt = (ctypes.c_double * 3)(0.5, 0.5, 0.0)
    assert N.lib().sph_probe_sum(g._h, ctypes.byref(p), 2, pt, 1, buf) == N.SPH_EINVAL
    assert N.lib().sph_probe_sum(g._h, ctypes.byref(p), 0, pt, -1, buf) == N.SPH_EINVAL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_probe_sums_equal_alone(precision):
    """In a batch, each reference system's probe sums (with its own mass)
    are bit-identical to the oracle's for that system; a periodic system is
    rejected."""
    rng = np.random.default_rng(8)
    b = sp.Batch(precision)
    cases = []
    for k, (n, h, m) in enumerate([(2000, 0.05, 0.3), (900, 0.08, 1.7)]):
        x = rng.random((n, 3)) * [1.0, 0.8, 0.0]
        v = rng.normal(size=(n, 3))
        v[:, 2] = 0.0
        typ = (rng.random(n) < 0.3).astype(float)
        sid = b.add(x, h, (0, 0, 0), (1.0, 0.8, 0.0), v=v, type=typ, **dict(PHYS, m=m))
        o = O.OracleSystem(h, np.zeros(3), np.array([1.0, 0.8, 0.0]))
        o.load(x, v, None, None, typ)
        assert o.create_cell_list() == n
        cases.append((sid, o, m, x))
    xp = rng.random((600, 3)) * [1.0, 1.0, 0.0]
    pid = b.add(xp, 0.1, (0, 0, 0), (1.0, 1.0, 0.0), periodic=True, **PHYS)
    b.build_cells()
    for sid, o, m, x in cases:
        P = probes(rng, x, 0.05, 2, np.array([1.0, 0.8, 1.0]))
        got = b.probe_sums(sid, P, fluid_only=True)
        assert np.array_equal(got, o.probe_sum(P, m, True), equal_nan=True)
    with pytest.raises(sp.SPHError):
        b.probe_sums(pid, [[0.5, 0.5, 0.0]])
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_probe_sums_aliased_grid(precision):
    """A grid one cell wide: three key_diff entries reach the probe's own
    cell, and rows wrap; the reference then visits cells several times and
    so must the device, in the same order."""
    rng = np.random.default_rng(12)
    n, h = 400, 0.1
    x = rng.random((n, 3)) * [0.05, 1.0, 0.0]
    v = rng.normal(size=(n, 3))
    typ = (rng.random(n) < 0.3).astype(float)
    hi = np.array([0.05, 1.0, 0.0])
    g, o = make_both(x, h, np.zeros(3), hi, precision, v, typ)
    assert sp.create_cell_list(g) == o.create_cell_list() == n
    P = rng.random((200, 3)) * [0.06, 1.0, 0.0]
    P[:20] = x[:20]
    got = sp.probe_sums(g, P, **PHYS)
    assert np.array_equal(got, o.probe_sum(P, PHYS["m"], False), equal_nan=True)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_particle_sum_lj_wall_bit_exact(dim, precision):
    """SmoothedParticles.sum(sys, LJ_potential, p) for every particle
    (src/core.jl:271-291; examples/collapse_symplectic.jl:144-162 with
    E_wall = 10 |g| H): bit-exact against the oracle on a dam break, before
    and after device steps.  The lattice puts the fluid one dr from the
    walls, so dr_wall = 1.5 dr here (the example's 0.95 dr only acts once the
    fluid presses into the walls)."""
    from smoothedparticles_hip import states
    st = states.dam_break(dim, 0.05 if dim == 2 else 0.1)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    lj = dict(m=st.m, dr_wall=1.5 * st.dr, E_wall=10 * 9.8 * 2.0, eps=1e-16)
    sp.init_reference_density(g, **p)
    for k in range(2):
        if k:
            sp.step(g, 30, scheme="wcsph_se_rho0", **p)
            sp.create_cell_list(g)
        o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
        o.load(g.get("x"), None, None, None, g.get("type"))
        o.create_cell_list()
        ref = o.particle_sum_lj(**lj)
        got = sp.particle_sums_lj_wall(g, **lj)
        assert np.array_equal(got, ref)
        assert np.count_nonzero(got) > 10 and np.all(got <= 0.0)
        assert np.all(got[st.type == 1.0] == 0.0) if k == 0 else True
