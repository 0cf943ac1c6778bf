"""Multi-step parity of the device time loop with the oracle.

FP64 dam break: the fused device step against the oracle's step; rounding
differences of the pair sums grow slowly over the steps, so the tolerance
is rel 1e-9 on rho, v and the total energy after 100 steps.
Mixed dam break: rho rel <= 1e-5 and total energy rel <= 1e-4 against the
FP64 oracle after 1000 steps (the stated mixed-precision bound).
"""
import numpy as np
import pytest

import oracle as O
import smoothedparticles_hip as sp
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu


def load(st, precision):
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
    o.load(st.x, st.v, st.rho, st.P, st.type)
    return g, o


def load_dambreak(st, precision):
    """Dam break with the per-particle reference density rho0_p = the
    initial density sum (tests/test_collision_2d.jl:123-124)."""
    g, o = load(st, precision)
    p = st.params()
    sp.init_reference_density(g, **p)
    o.create_cell_list()
    o.density_sum_rho0(True, **p)
    tol = 1e-12 if precision == "fp64" else 5e-6
    assert np.max(np.abs(g.get("rho0") / o.field("rho0") - 1)) <= tol
    return g, o


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("scheme", ["wcsph_se", "wcsph_se_rho0"])
def test_fused_step_equals_operator_sequence(scheme):
    """The fused step gives the bits of the operators applied one by one,
    for the constant-rho0 and the rho0-field equation of state."""
    st = states.dam_break(2, 0.02)
    p = st.params()
    g1, _ = load(st, "fp64")
    g2, _ = load(st, "fp64")
    eos = sp.find_pressure if scheme == "wcsph_se" else sp.find_pressure_rho0
    if scheme == "wcsph_se_rho0":
        for g in (g1, g2):
            sp.init_reference_density(g, **p)
    sp.step(g1, 5, scheme=scheme, **p)
    for _ in range(5):
        sp.create_cell_list(g2)
        sp.apply(g2, sp.reset_rho)
        sp.apply(g2, sp.find_rho, self=True, **p)
        sp.apply(g2, eos, **p)
        sp.apply(g2, sp.reset_a)
        sp.apply(g2, sp.internal_force, **p)
        sp.apply(g2, sp.accelerate, **p)
        a_last = g2.get("a")
        sp.apply(g2, sp.move, **p)   # resets a (collapse_dry.jl:149), the fused step keeps it
        assert np.all(g2.get("a") == 0.0)
    for f in ("x", "v", "rho", "P"):
        assert np.array_equal(g1.get(f), g2.get(f)), f
    assert np.array_equal(g1.get("a"), a_last)


def test_dambreak_2d_fp64_100_steps():
    st = states.dam_break(2, 0.02)
    p = st.params()
    g, o = load_dambreak(st, "fp64")
    pairs_g = sp.step(g, 100, scheme="wcsph_se_rho0", **p)
    pairs_o = o.step_wcsph(100, use_rho0=True, **p)
    assert pairs_g == pairs_o
    assert len(g) == o.n == len(st.x)
    assert rel(g.get("rho"), o.field("rho")) <= 1e-9
    assert rel(g.get("v"), o.field("v")) <= 1e-9
    assert sp.energy(g, "wcsph_rho0", **p) == pytest.approx(o.energy("wcsph_rho0", **p), rel=1e-9)


def test_dambreak_2d_mixed_1000_steps():
    """Mixed against the FP64 oracle at steps 100 and 1000: rho rel <= 1e-5,
    total energy rel <= 1e-4.  (Positions drift apart slowly, so the pair
    counts differ by a few 1e-4 after 1000 steps; they are not compared.)"""
    st = states.dam_break(2, 0.02)
    p = st.params()
    g, o = load_dambreak(st, "mixed")
    for steps in (100, 900):
        sp.step(g, steps, scheme="wcsph_se_rho0", **p)
        o.step_wcsph(steps, use_rho0=True, **p)
        assert len(g) == o.n == len(st.x)
        err_rho = np.max(np.abs(g.get("rho") / o.field("rho") - 1))
        err_e = abs(sp.energy(g, "wcsph_rho0", **p) / o.energy("wcsph_rho0", **p) - 1)
        print(f"after {steps}: rho rel {err_rho:.3e}, E rel {err_e:.3e}")
        assert err_rho <= 1e-5
        assert err_e <= 1e-4


def test_dambreak_3d_fp64_small():
    st = states.dam_break(3, 0.05)
    p = st.params()
    g, o = load_dambreak(st, "fp64")
    assert sp.step(g, 20, scheme="wcsph_se_rho0", **p) == o.step_wcsph(20, use_rho0=True, **p)
    assert rel(g.get("rho"), o.field("rho")) <= 1e-10
    assert rel(g.get("x"), o.field("x")) <= 1e-12


def test_collision_2d_energy_on_device():
    """tests/test_collision_2d.jl on the device: count constant and energy
    drift < 1e-2, and the energy trace follows the oracle's: rel 1e-9 before
    the discs touch (first two samples) and rel 1e-3 over the whole trace
    (rounding differences of the pair sums grow once the collision turns
    chaotic)."""
    st = states.collision_2d()
    kw = dict(m=st.m, rho0=st.rho0, c=st.c, dt=st.dt)
    g, o = load(st, "fp64")
    sp.create_cell_list(g)
    sp.apply(g, sp.find_rho0, self=True, **kw)
    sp.apply(g, sp.find_rho, self=True, **kw)
    sp.apply(g, sp.find_pressure_rho0, **kw)
    sp.apply(g, sp.internal_force_rho0, **kw)
    steps = int(round(1.0 / st.dt))
    every = int(round(0.1 / st.dt))
    N, E = [], []
    ops = sp.program_operators("test_collision_2d")
    for k in range(steps + 1):
        # verlet_step! (tests/test_collision_2d.jl:106-116), by the program's own names
        sp.apply(g, ops["accelerate!"], **kw)
        sp.apply(g, ops["move!"], **kw)
        sp.create_cell_list(g)
        sp.apply(g, ops["reset_rho!"])
        sp.apply(g, ops["find_rho!"], self=True, **kw)
        sp.apply(g, ops["find_pressure!"], **kw)
        sp.apply(g, ops["reset_a!"])
        sp.apply(g, ops["internal_force!"], **kw)
        sp.apply(g, ops["accelerate!"], **kw)
        if k % every == 0:
            N.append(len(g))
            E.append(sp.energy(g, "collision", **kw))
    assert all(n == N[0] for n in N)
    E = np.array(E)
    assert (E / E[0] - 1).max() < 1e-2
    from test_oracle_collision import run_collision
    _, Eo = run_collision(o, st, steps, every)
    d = np.abs(E / np.array(Eo) - 1)
    print("energy trace rel diff vs oracle:", " ".join(f"{v:.1e}" for v in d))
    assert d[:2].max() <= 1e-9
    assert d.max() <= 1e-3


@pytest.mark.parametrize("dr", [0.0125])
def test_full_size_3d_step0_mixed_vs_fp64(dr):
    """At the 1M-fluid 3D dam break: the mixed neighbour count equals the
    FP64 one and the step-0 densities agree to 5e-6."""
    st = states.dam_break(3, dr)
    g64 = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "fp64")
    g32 = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "mixed")
    for g in (g64, g32):
        g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    p = st.params()
    for g in (g64, g32):
        sp.init_reference_density(g, **p)
    assert np.array_equal(g64.get("rho0") > 0, np.ones(len(g64), bool))
    n64 = sp.step(g64, 1, scheme="wcsph_se_rho0", **p)
    n32 = sp.step(g32, 1, scheme="wcsph_se_rho0", **p)
    assert n64 == n32
    assert n64 % 2 == 0  # the neighbour relation is symmetric
    r64, r32 = g64.get("rho"), g32.get("rho")
    assert np.max(np.abs(r32 / r64 - 1)) <= 5e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_density_pass_distance_tests(precision):
    """The density pass's distance tests (sph_step_tests): every accepted
    pair was tested, and the bounding-box cull tests fewer candidates than
    the reference's 27-cell search visits."""
    st = states.dam_break(3, 0.04)
    p = st.params()
    g, _ = load_dambreak(st, precision)
    sp.step(g, 1, scheme="wcsph_se_rho0", **p)
    dens_pairs, _ = g.step_pairs()
    tests = g.step_tests()
    sp.create_cell_list(g)  # positions moved by one step; the lists barely change
    cand_all, _ = g.candidates()
    assert 0 < dens_pairs <= tests
    assert tests < cand_all
    g.close()
