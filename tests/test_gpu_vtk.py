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


def test_frame_writer_matches_save_frame(tmp_path):
    """Frames written through device snapshots while the steps go on are the
    files save_frame writes when the loop stops for each frame."""
    import time
    st = states.dam_break(3, 0.02)
    p = st.params()
    gs = []
    for _ in range(2):
        g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "mixed")
        g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
        g.set_tile(64)
        sp.init_reference_density(g, **p)
        gs.append(g)
    ds = V.new_pvd_file(str(tmp_path / "sync"))
    da = V.new_pvd_file(str(tmp_path / "async"))
    w = V.FrameWriter(da)
    t = {}
    for label, g in (("sync", gs[0]), ("async", gs[1])):
        sp.step(g, 2, scheme="wcsph_se_rho0", **p)
        t0 = time.perf_counter()
        for k in range(4):
            if label == "sync":
                V.save_frame(ds, g, "v", "rho", "P", "type")
            else:
                w.save(g, "v", "rho", "P", "type")
            sp.step(g, 10, scheme="wcsph_se_rho0", **p)
        if label == "async":
            w.flush()
        t[label] = time.perf_counter() - t0
    w.close()
    V.save_pvd_file(ds)
    V.save_pvd_file(da)
    print(f"{len(st.x)} particles, 4 frames + 40 steps: sync {t['sync']:.3f} s, "
          f"async {t['async']:.3f} s")
    for k in range(4):
        a = (tmp_path / "sync" / f"frame{k}.vtp").read_bytes()
        b = (tmp_path / "async" / f"frame{k}.vtp").read_bytes()
        assert a == b, k
    assert (tmp_path / "sync" / "result.pvd").read_bytes() == (tmp_path / "async" / "result.pvd").read_bytes()


def test_snapshot_is_the_state_at_the_call():
    st = states.dam_break(2, 0.02)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "fp64")
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    sp.init_reference_density(g, **p)
    sp.step(g, 5, scheme="wcsph_se_rho0", **p)
    ref = {k: g.get(k) for k in ("x", "v", "rho")}
    with g.snapshot("x", "v", "rho") as snap:
        sp.step(g, 7, scheme="wcsph_se_rho0", **p)   # later steps do not reach the snapshot
        got = snap.wait()
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
    assert not np.array_equal(g.get("x"), ref["x"])
    b = sp.Batch("mixed")
    tg = states.taylor_green(30, 100.0)
    sid = b.add_state(tg, periodic=True)
    b.init_rho0()
    b.step(3)
    with b.snapshot(sid, "v") as snap:
        v = b.get(sid, "v")
        assert np.array_equal(snap.wait()["v"], v)
    with pytest.raises(sp.SPHError):
        g.snapshot("x", "x", "x", "x", "x", "x", "x", "x", "x", "x")
    b.close()
