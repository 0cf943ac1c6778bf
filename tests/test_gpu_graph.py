"""Steps replayed from a hipGraph (SPH_OPTION_GRAPH) give the bits of the
same steps enqueued kernel by kernel.  Each graph is captured once per step
call and replayed several times; the calls are separated by the program's
own eager work (field reads, reductions, a rebuild), and every call's result
is compared with the eager run's."""
import time

import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "a", "rho", "P", "Drho")


def make(st, precision, graph):
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    g.set_graph(graph)
    return g


@pytest.mark.parametrize("scheme", ["wcsph_se_rho0", "collapse_dry"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_steps_equal_eager(scheme, precision):
    st = states.dam_break(2, 0.04)
    p = st.params()
    if scheme == "collapse_dry":
        p.update(mu=8.4e-4, nu=1e-6, alpha=0.0)
    gs = [make(st, precision, False), make(st, precision, True)]
    for g in gs:
        if scheme == "collapse_dry":
            sp.collapse_dry_init(g, **p)
        else:
            sp.init_reference_density(g, **p)
    for n in (7, 10, 4, 1, 6):
        res = []
        for g in gs:
            pairs = sp.step(g, n, scheme=scheme, **p)
            res.append((pairs, g.step_pairs(), {f: g.get(f) for f in FIELDS},
                        sp.energy(g, "wcsph", **p)))
        assert res[0][0] == res[1][0] and res[0][1] == res[1][1], n
        for f in FIELDS:
            assert np.array_equal(res[0][2][f], res[1][2][f]), (n, f)
        assert res[0][3] == res[1][3]
        # eager work between the calls: a rebuild and a probe
        for g in gs:
            sp.create_cell_list(g)
        assert np.array_equal(sp.probe_sums(gs[0], st.x[:20], **p), sp.probe_sums(gs[1], st.x[:20], **p))


def test_graph_batch_equal_eager():
    sts = [(states.taylor_green(40, 200.0), True), (states.dam_break(2, 0.05), False),
           (states.taylor_green(56, 600.0), True)]
    out = []
    for graph in (False, True):
        b = sp.Batch("mixed")
        b.set_tile(32)
        b.set_graph(graph)
        ids = [b.add_state(s, periodic=per) for s, per in sts]
        b.init_rho0()
        snaps = []
        for n in (9, 12, 5):
            b.step(n)
            snaps.append([b.get(i, "v") for i in ids])
        out.append(snaps)
        b.close()
    for a, c in zip(out[0], out[1]):
        for va, vc in zip(a, c):
            assert np.array_equal(va, vc)


def test_graph_launch_overhead_small_system():
    """A served-size request (about 1.5k particles, 2D): time per step
    enqueued kernel by kernel and replayed from the graph."""
    st = states.dam_break(2, 0.05)
    p = st.params()
    t = {}
    for graph in (False, True):
        g = make(st, "mixed", graph)
        sp.init_reference_density(g, **p)
        sp.step(g, 20, scheme="wcsph_se_rho0", **p)   # warm-up
        t0 = time.perf_counter()
        sp.step(g, 400, scheme="wcsph_se_rho0", **p)
        t[graph] = (time.perf_counter() - t0) / 400
    print(f"{len(st.x)} particles: eager {t[False] * 1e6:.1f} us/step, "
          f"graph {t[True] * 1e6:.1f} us/step")
    assert t[True] < t[False] * 1.1
