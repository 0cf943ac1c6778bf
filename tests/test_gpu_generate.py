"""Device-side initial-state generation (sph_generate_lattice,
include/sph_setup.h): generate_particles! of the dam-break fluid block and
wall layer written straight into the device arrays gives the bits of the
host-generated state (positions and their order, jitter, hydrostatic rho and
P, types), up to the full 3D benchmark size."""
import time

import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,dr", [(2, 0.02), (2, 0.013), (3, 0.04), (3, 0.0079)])
def test_device_dam_break_equals_host(dim, dr):
    t0 = time.perf_counter()
    st = states.dam_break(dim, dr)
    g1 = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "mixed")
    g1.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    t_host = time.perf_counter() - t0
    t0 = time.perf_counter()
    geo, fs, ws = states.dam_break_specs(dim, dr)
    g2 = sp.ParticleSystem(sp.Box(*geo.bbox_lo, *geo.bbox_hi), geo.R, "mixed")
    nf = g2.generate_lattice(fs)
    nw = g2.generate_lattice(ws)
    t_dev = time.perf_counter() - t0
    print(f"dim {dim} dr {dr}: {nf} fluid + {nw} wall; host state + upload {t_host:.3f} s, "
          f"device {t_dev:.3f} s")
    assert nf == st.n_fluid and nw == st.n_wall
    assert np.array_equal(geo.bbox_lo, st.bbox_lo) and np.array_equal(geo.bbox_hi, st.bbox_hi)
    for f in ("x", "v", "rho", "P", "type", "a", "rho0"):
        assert np.array_equal(g2.get(f), g1.get(f)), f
    if dr >= 0.013:
        p = st.params()
        for g in (g1, g2):
            sp.init_reference_density(g, **p)
            sp.step(g, 3, scheme="wcsph_se_rho0", **p)
        assert np.array_equal(g1.get("x"), g2.get("x"))


def test_generate_into_a_batch_and_errors():
    geo, fs, ws = states.dam_break_specs(2, 0.04)
    st = states.dam_break(2, 0.04)
    b = sp.Batch("fp64")
    tg = states.taylor_green(20, 100.0)
    b.add_state(tg, periodic=True)
    sid = b.add(np.zeros((0, 3)), geo.R, geo.bbox_lo, geo.bbox_hi, **st.params())
    assert b.generate_lattice(sid, fs) == st.n_fluid
    assert b.generate_lattice(sid, ws) == st.n_wall
    assert np.array_equal(b.get(sid, "x"), st.x) and np.array_equal(b.get(sid, "P"), st.P)
    with pytest.raises(sp.SPHError):   # particles go to the last system only
        b.generate_lattice(0, fs)
    bad = states.lattice_spec(5, 2, 0.04, (0, 0, 0), (1, 1, 0))
    with pytest.raises(sp.SPHError):
        b.generate_lattice(sid, bad)
    b.close()
