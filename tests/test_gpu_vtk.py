"""Frames from the device and resume from a frame (src/IO.jl:53-121): the
fields downloaded for a frame equal the system's, and a system imported from
a frame continues exactly as the original does."""
import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import states, vtk_io as V

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_frame_resume_is_exact(tmp_path, precision):
    st = states.dam_break(2, 0.01)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    sp.init_reference_density(g, **p)
    sp.step(g, 20, scheme="wcsph_se_rho0", **p)
    d = V.new_pvd_file(str(tmp_path / "run"))
    V.save_frame(d, g, "v", "a", "rho", "P", "Drho", "rho0", "type")
    V.save_pvd_file(d)
    (_, path), = V.read_pvd(str(tmp_path / "run" / "result.pvd"))
    h = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    assert V.import_particles(h, path) == len(g)
    for k in ("x", "v", "a", "rho", "P", "Drho", "rho0", "type"):
        assert np.array_equal(h.get(k), g.get(k)), k
    g.set_tile(32)  # the per-run partial sums depend on the tile size
    h.set_tile(32)
    sp.step(g, 10, scheme="wcsph_se_rho0", **p)
    sp.step(h, 10, scheme="wcsph_se_rho0", **p)
    for k in ("x", "v", "rho"):
        assert np.array_equal(h.get(k), g.get(k)), k
