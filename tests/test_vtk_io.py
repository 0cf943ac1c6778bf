"""Frame output and particle import (src/IO.jl:9-121): the VTK XML PolyData
writer and reader and the ParaView collection.  CPU only; the device side
(downloading fields for a frame, resuming from one) is tested in
test_gpu_vtk.py.

Pinned by the reference's own data file examples/init/cylinder.vtp (written
by WriteVTK, read by examples/cylinder.jl:85): reading it gives the particle
lattice and types, and writing those arrays back reproduces the file byte for
byte."""
import os

import numpy as np
import pytest

from smoothedparticles_hip import vtk_io as V

CYL = "/root/reference/examples/init/cylinder.vtp"


@pytest.mark.skipif(not os.path.exists(CYL), reason="reference data file not present")
def test_reference_file_reads_and_rewrites_byte_for_byte(tmp_path):
    pts, f = V.read_vtp(CYL)
    n = 16912
    assert pts.shape == (n, 3)
    assert np.all(pts[:, 2] == 0.0)
    assert np.array_equal(f["__connectivity__"], np.arange(n))
    assert np.array_equal(f["__offsets__"], np.arange(1, n + 1))
    vals, counts = np.unique(f["type"], return_counts=True)
    assert list(vals) == [0.0, 1.0, 2.0, 3.0]
    assert counts.sum() == n
    out = tmp_path / "c.vtp"
    V.write_vtp(out, pts, [("type", f["type"])])
    assert out.read_bytes() == open(CYL, "rb").read()


@pytest.mark.parametrize("compress", [True, False])
def test_round_trip(tmp_path, compress):
    rng = np.random.default_rng(0)
    n = 1000
    pts = rng.normal(size=(n, 3))
    v = rng.normal(size=(n, 3))
    rho = 1000 + rng.random(n)
    typ = (rng.random(n) < 0.3).astype(float)
    pts[5] = [np.nan, np.inf, -0.0]
    p = tmp_path / "f.vtp"
    V.write_vtp(p, pts, [("v", v), ("rho", rho), ("type", typ)], compress=compress)
    q, f = V.read_vtp(p)
    assert np.array_equal(q, pts, equal_nan=True)
    assert np.signbit(q[5, 2])
    assert np.array_equal(f["v"], v) and np.array_equal(f["rho"], rho)
    assert np.array_equal(f["type"], typ)


def test_empty_frame_and_ascii_arrays(tmp_path):
    p = tmp_path / "e.vtp"
    V.write_vtp(p, np.zeros((0, 3)), [("rho", np.zeros(0))])
    q, f = V.read_vtp(p)
    assert q.shape == (0, 3) and f["rho"].shape == (0,)
    a = tmp_path / "a.vtp"
    a.write_text('<?xml version="1.0"?>\n<VTKFile type="PolyData" version="1.0" '
                 'byte_order="LittleEndian"><PolyData><Piece NumberOfPoints="2" NumberOfVerts="0">'
                 '<Points><DataArray type="Float64" NumberOfComponents="3" format="ascii">'
                 '0 1 2 3 4 5</DataArray></Points><PointData><DataArray type="Float64" '
                 'Name="P" format="ascii">7.5 -1</DataArray></PointData></Piece></PolyData>'
                 '</VTKFile>')
    q, f = V.read_vtp(a)
    assert np.array_equal(q, [[0, 1, 2], [3, 4, 5]]) and np.array_equal(f["P"], [7.5, -1])


class HostSystem:
    """Host stand-in with the ParticleSystem field interface."""

    def __init__(self):
        self.f = {k: np.zeros((0, 3) if k in ("x", "v", "a") else (0,))
                  for k in ("x", "v", "a", "rho", "P", "Drho", "rho0", "type")}

    def __len__(self):
        return len(self.f["x"])

    def get(self, name):
        return self.f[name].copy()

    def set(self, name, a):
        self.f[name] = np.array(a, dtype=float).reshape(self.f[name].shape)

    def add_particles(self, x, v=None, rho=None, P=None, type=None):
        n = len(x)
        for k, a in (("x", x), ("v", v), ("rho", rho), ("P", P), ("type", type)):
            shape = (n, 3) if k in ("x", "v") else (n,)
            new = np.zeros(shape) if a is None else np.broadcast_to(a, shape)
            self.f[k] = np.concatenate([self.f[k], new])
        for k in ("a", "Drho", "rho0"):
            shape = (n, 3) if k == "a" else (n,)
            self.f[k] = np.concatenate([self.f[k], np.zeros(shape)])


def test_frames_collection_and_import(tmp_path):
    rng = np.random.default_rng(1)
    s = HostSystem()
    s.add_particles(rng.random((50, 3)), v=rng.random((50, 3)), rho=rng.random(50),
                    type=np.ones(50))
    s.set("rho0", rng.random(50))
    d = V.new_pvd_file(str(tmp_path / "out" / "run"))
    for _ in range(3):
        V.save_frame(d, s, "v", "rho", "type", "rho0")
    V.save_pvd_file(d)
    frames = V.read_pvd(str(tmp_path / "out" / "run" / "result.pvd"))
    assert [t for t, _ in frames] == [0.0, 1.0, 2.0]
    assert all(os.path.basename(p) == f"frame{k}.vtp" for k, (_, p) in enumerate(frames))
    t = HostSystem()
    t.add_particles(np.zeros((4, 3)))
    assert V.import_particles(t, frames[-1][1], P=3.0) == 50
    assert len(t) == 54
    for k in ("x", "v", "rho", "type", "rho0"):
        assert np.array_equal(t.get(k)[4:], s.get(k))
    assert np.all(t.get("P")[4:] == 3.0) and np.all(t.get("P")[:4] == 0.0)
    with pytest.raises(KeyError):
        V.save_frame(d, s, "pressure")
