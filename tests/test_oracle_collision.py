"""tests/test_collision_2d.jl restated on the oracle: two colliding discs,
particle count constant and max relative energy drift < 1e-2
(tests/test_collision_2d.jl:141-147).  CPU only."""
import numpy as np

import oracle as O
from smoothedparticles_hip import states


def run_collision(o, st, steps, every):
    kw = dict(m=st.m, rho0=st.rho0, c=st.c, dt=st.dt)
    # initialisation (tests/test_collision_2d.jl:123-127)
    o.create_cell_list()
    o.density_sum_rho0(True, **kw)
    o.density_sum(True, **kw)
    o.eos(use_field=True, **kw)
    o.pressure_force_rho0(**kw)
    N, E = [], []
    for k in range(steps + 1):
        # verlet_step! (tests/test_collision_2d.jl:106-116)
        o.kick(coeff=0.5, **kw)      # accelerate!, :92-94
        o.drift(coeff=1.0, **kw)     # move!, :88-90
        o.create_cell_list()
        o.reset_rho()
        o.density_sum(True, **kw)
        o.eos(use_field=True, **kw)
        o.reset_a()
        o.pressure_force_rho0(**kw)
        o.kick(coeff=0.5, **kw)
        if k % every == 0:
            N.append(o.n)
            E.append(o.energy("collision", **kw))
    return N, E


def test_collision_energy_conservation():
    st = states.collision_2d()
    o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
    o.load(st.x, st.v, st.rho, st.P, st.type)
    steps = int(round(1.0 / st.dt))
    N, E = run_collision(o, st, steps, int(round(0.1 / st.dt)))
    assert all(n == N[0] for n in N)
    E = np.array(E)
    assert (E / E[0] - 1.0).max() < 1e-2
