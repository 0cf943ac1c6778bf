"""ParaView frame output and particle import (src/IO.jl:9-121).

The reference writes each frame with WriteVTK as a VTK XML PolyData file
(`.vtp`: the particle positions as points, one vertex cell per particle, the
chosen fields as point data, appended raw binary, zlib-compressed, UInt64
block headers) and lists the frames in a ParaView collection (`result.pvd`)
with the frame number as the time step.  `import_particles` reads such a
file back into a system, which is how the reference resumes a run
(examples/cylinder.jl:85).

save_frame downloads the fields synchronously (one copy per field, in the
reference's particle order).  FrameWriter does the same without stopping the
steps: it starts an asynchronous device snapshot (sph_snapshot_begin: the
fields are gathered on the system's stream, so they are the state of the
call, and copied to pinned memory on a second stream while the next steps
run) and encodes and writes the file on a worker thread.
"""
from concurrent.futures import ThreadPoolExecutor
import os
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

from . import _native as N

_VTK_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64,
              "Int32": np.int32, "UInt64": np.uint64, "UInt32": np.uint32, "UInt8": np.uint8}
_NP_TO_VTK = {np.dtype(v).str: k for k, v in _VTK_TYPES.items()}


# ------------------------------------------------------------- encoding --
def _block(arr, compress):
    """One appended data block: UInt64 header then the bytes.  Compressed
    arrays are one zlib block, header [nblocks, block size, last block size,
    compressed size] as WriteVTK writes them."""
    raw = np.ascontiguousarray(arr).astype(np.dtype(arr.dtype).newbyteorder("<")).tobytes()
    if not compress:
        return struct.pack("<Q", len(raw)) + raw
    if not raw:
        return struct.pack("<3Q", 0, 0, 0)
    c = zlib.compress(raw)
    return struct.pack("<4Q", 1, len(raw), len(raw), len(c)) + c


def write_vtp(path, points, fields=(), compress=True):
    """Write a PolyData file: points (n, 3) and fields, a sequence of
    (name, array) with arrays of shape (n,) or (n, k)."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    n = pts.shape[0]
    arrays = [("Points", pts, "Points"),
              ("connectivity", np.arange(n, dtype=np.int64), "Verts"),
              ("offsets", np.arange(1, n + 1, dtype=np.int64), "Verts")]
    for name, a in fields:
        a = np.asarray(a)
        if a.shape[0] != n:
            raise ValueError(f"field {name!r} has {a.shape[0]} values for {n} points")
        arrays.append((str(name), np.ascontiguousarray(a), "PointData"))
    blobs, offs, off = [], [], 0
    for _, a, _ in arrays:
        b = _block(a, compress)
        blobs.append(b)
        offs.append(off)
        off += len(b)

    def tag(i, indent):
        name, a, _ = arrays[i]
        ncomp = 1 if a.ndim == 1 else a.shape[1]
        vtype = _NP_TO_VTK[np.dtype(a.dtype).newbyteorder("<").str]
        return (f'{indent}<DataArray type="{vtype}" Name="{name}" NumberOfComponents="{ncomp}" '
                f'format="appended" offset="{offs[i]}"/>\n')

    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    head = ('<?xml version="1.0" encoding="utf-8"?>\n'
            f'<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
            f'header_type="UInt64"{comp}>\n'
            '  <PolyData>\n'
            f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}">\n'
            '      <Points>\n' + tag(0, "        ") + '      </Points>\n'
            '      <Verts>\n' + tag(1, "        ") + tag(2, "        ") + '      </Verts>\n'
            '      <PointData>\n'
            + "".join(tag(i, "        ") for i in range(3, len(arrays))) +
            '      </PointData>\n'
            '    </Piece>\n'
            '  </PolyData>\n'
            '  <AppendedData encoding="raw">\n_')
    with open(path, "wb") as f:
        f.write(head.encode())
        for b in blobs:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>")


# ------------------------------------------------------------- decoding --
def _decode(buf, pos, dtype, count, header, compressed):
    hs = struct.calcsize("<" + header)
    if not compressed:
        (nbytes,) = struct.unpack_from("<" + header, buf, pos)
        raw = bytes(buf[pos + hs: pos + hs + nbytes])
    else:
        nb, bs, lbs = struct.unpack_from("<3" + header, buf, pos)
        sizes = struct.unpack_from(f"<{nb}{header}", buf, pos + 3 * hs)
        p = pos + (3 + nb) * hs
        parts = []
        for k, cs in enumerate(sizes):
            part = zlib.decompress(bytes(buf[p:p + cs]))
            want = lbs if (k == nb - 1 and lbs != 0) else bs
            if len(part) != want:
                raise ValueError("VTK block size mismatch")
            parts.append(part)
            p += cs
        raw = b"".join(parts)
    a = np.frombuffer(raw, dtype=np.dtype(dtype).newbyteorder("<"))
    if a.size != count:
        raise ValueError(f"VTK array holds {a.size} values, expected {count}")
    return a.astype(dtype)


def read_vtp(path):
    """Read a PolyData file with appended raw (optionally zlib-compressed)
    or ASCII arrays.  Returns (points (n, 3), {name: array}) where point
    data arrays have shape (n,) or (n, k)."""
    buf = open(path, "rb").read()
    i = buf.find(b"<AppendedData")
    if i >= 0:
        start = buf.find(b"_", buf.find(b">", i)) + 1
        root = ET.fromstring(buf[:i].decode("utf-8") + "</VTKFile>")
    else:
        start = -1
        root = ET.fromstring(buf.decode("utf-8"))
    if root.get("type") != "PolyData":
        raise ValueError(f"{path}: not a PolyData file")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError(f"{path}: big-endian VTK files are not supported")
    header = "Q" if root.get("header_type", "UInt32") == "UInt64" else "I"
    compressed = root.get("compressor") is not None
    if compressed and root.get("compressor") != "vtkZLibDataCompressor":
        raise ValueError(f"{path}: unsupported compressor {root.get('compressor')}")
    piece = root.find("PolyData/Piece")
    n = int(piece.get("NumberOfPoints"))

    def load(el, nrow):
        dtype = _VTK_TYPES[el.get("type")]
        ncomp = int(el.get("NumberOfComponents", "1"))
        fmt = el.get("format")
        if fmt == "appended":
            a = _decode(buf, start + int(el.get("offset")), dtype, nrow * ncomp, header, compressed)
        elif fmt == "ascii":
            a = np.array(el.text.split(), dtype=dtype)
        else:
            raise ValueError(f"{path}: DataArray format {fmt!r} is not supported")
        return a.reshape(nrow, ncomp) if ncomp > 1 else a.reshape(nrow)

    points = load(piece.find("Points/DataArray"), n).astype(np.float64).reshape(n, 3)
    fields = {}
    for el in piece.findall("PointData/DataArray"):
        fields[el.get("Name")] = load(el, n)
    verts = {el.get("Name"): el for el in piece.findall("Verts/DataArray")}
    nv = int(piece.get("NumberOfVerts", "0"))
    if "connectivity" in verts and "offsets" in verts and nv:
        offs = load(verts["offsets"], nv)
        conn = load(verts["connectivity"], int(offs[-1]))
        fields["__connectivity__"], fields["__offsets__"] = conn, offs
    return points, fields


# ------------------------------------------------------ DataStorage API --
class DataStorage:
    """DataStorage (src/IO.jl:9-13): the output folder, the frames written
    so far and the next frame number."""

    def __init__(self, path):
        self.path = path
        self.frame = 0
        self.frames = []  # (time step, file name relative to path)


def new_pvd_file(path):
    """new_pvd_file (src/IO.jl:20-26): creates the folder if needed."""
    os.makedirs(path, exist_ok=True)
    return DataStorage(path)


def save_frame(data, sys, *vars, compress=True):
    """save_frame! (src/IO.jl:53-75): writes frame<k>.vtp with the particle
    positions and the named fields, and records it under time step k."""
    pts = sys.get("x")
    fields = []
    for var in vars:
        if var not in N.FIELDS:
            raise KeyError(f"Variable {var} does not exist!")
        fields.append((var, sys.get(var)))
    name = f"frame{data.frame}.vtp"
    write_vtp(os.path.join(data.path, name), pts, fields, compress=compress)
    data.frames.append((data.frame, name))
    data.frame += 1


class FrameWriter:
    """save_frame! at frame cadence without blocking the time loop.

    w = FrameWriter(data); w.save(sys, "v", "rho") starts the frame's device
    snapshot and returns at once; the frame gets the next frame number of
    `data` now, and its file is written by a worker thread once the copy is
    done.  flush() waits for every pending frame; close() also stops the
    workers.  Frames are the same files save_frame writes."""

    def __init__(self, data, workers=2, compress=True):
        self.data = data
        self.compress = compress
        self.pool = ThreadPoolExecutor(max_workers=workers)
        self.pending = []

    def _write(self, snap, path, vars):
        try:
            arrays = snap.wait()
        finally:
            snap.close()
        write_vtp(path, arrays["x"], [(v, arrays[v]) for v in vars], compress=self.compress)

    def save(self, sys, *vars, sid=None):
        for var in vars:
            if var not in N.FIELDS:
                raise KeyError(f"Variable {var} does not exist!")
        names = ("x",) + tuple(v for v in vars if v != "x")
        snap = sys.snapshot(*names) if sid is None else sys.snapshot(sid, *names)
        name = f"frame{self.data.frame}.vtp"
        self.data.frames.append((self.data.frame, name))
        self.data.frame += 1
        self.pending.append(self.pool.submit(self._write, snap, os.path.join(self.data.path, name),
                                             vars))

    def flush(self):
        for f in self.pending:
            f.result()
        self.pending = []

    def close(self):
        self.flush()
        self.pool.shutdown()


def save_pvd_file(data):
    """save_pvd_file (src/IO.jl:33-35): writes the collection result.pvd."""
    lines = ['<?xml version="1.0" encoding="utf-8"?>',
             '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">',
             '  <Collection>']
    for t, name in data.frames:
        lines.append(f'    <DataSet timestep="{t}" part="0" file="{name}"/>')
    lines += ['  </Collection>', '</VTKFile>', '']
    with open(os.path.join(data.path, "result.pvd"), "w") as f:
        f.write("\n".join(lines))


def read_pvd(path):
    """The (time step, file path) entries of a ParaView collection."""
    root = ET.parse(path).getroot()
    base = os.path.dirname(path)
    return [(float(d.get("timestep")), os.path.join(base, d.get("file")))
            for d in root.findall("Collection/DataSet")]


def import_particles(sys, path, **defaults):
    """import_particles! (src/IO.jl:83-121): appends the file's points as
    particles and sets every field of the particle type that the file holds.
    The reference builds each particle with a constructor before copying the
    fields; `defaults` stand in for it (field name -> value for all imported
    particles, e.g. type=1.0), and the file's fields override them.
    Returns the number of particles imported."""
    pts, fields = read_vtp(path)
    n = pts.shape[0]
    vals = {}
    for name in N.FIELDS:
        if name == "x":
            continue
        if name in fields:
            a = np.asarray(fields[name], dtype=np.float64)
            want = (n, 3) if name in N.VECTOR_FIELDS else (n,)
            if a.shape != want:
                raise ValueError(f"{path}: field {name!r} has shape {a.shape}, expected {want}")
            vals[name] = a
        elif name in defaults:
            vals[name] = defaults[name]
    n0 = len(sys)
    sys.add_particles(pts, v=vals.get("v"), rho=vals.get("rho"), P=vals.get("P"),
                      type=vals.get("type"))
    for name in ("a", "Drho", "rho0"):
        if name in vals:
            cur = sys.get(name)
            cur[n0:] = vals[name]
            sys.set(name, cur)
    return n
