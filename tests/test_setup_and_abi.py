"""Host-side checks that need no GPU: the lattice generator against an
independent numpy restatement of covering / BoundaryLayer
(src/grids.jl:52-66,128-144, src/geometry.jl:198-234), the C-ABI symbol
table of libsph_hip, and the operator registry of the Python mirror."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import _native as N
from smoothedparticles_hip import states

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def np_covering_box(dim, dr, lo, hi):
    r = [range(math.floor(lo[d] / dr), math.ceil(hi[d] / dr) + 1) for d in range(dim)]
    if dim == 2:
        r.append(range(0, 1))
    I, J, K = np.meshgrid(np.array(r[0], float), np.array(r[1], float), np.array(r[2], float),
                          indexing="ij")
    x = np.stack([I * dr, J * dr, (K * dr if dim == 3 else np.zeros_like(K))], -1).reshape(-1, 3)
    inside = np.all((x >= np.asarray(lo)) & (x <= np.asarray(hi)), axis=1)
    return x[inside]


def np_walls(dim, dr, tlo, thi, w, lid_axis, lid):
    tlo, thi = np.asarray(tlo, float), np.asarray(thi, float)
    blo, bhi = tlo.copy(), thi.copy()
    blo[:dim] -= w
    bhi[:dim] += w
    ball = np_covering_box(dim, dr, -np.full(3, w) * [1, 1, 1 if dim == 3 else 0],
                           np.full(3, w) * [1, 1, 1 if dim == 3 else 0])
    ball = ball[((ball[:, 0] - 0.0) ** 2 + (ball[:, 1] - 0.0) ** 2) + (ball[:, 2] - 0.0) ** 2 <= w * w]
    cand = np_covering_box(dim, dr, blo, bhi)
    cand = cand[cand[:, lid_axis] < lid]
    in_tank = np.all((cand >= tlo) & (cand <= thi), axis=1)
    cand = cand[~in_tank]
    hit = np.zeros(len(cand), bool)
    for dx in ball:
        y = cand + dx
        hit |= np.all((y >= tlo) & (y <= thi), axis=1)
    return cand[hit]


@pytest.mark.parametrize("dim,dr", [(2, 0.02), (3, 0.1)])
def test_lattice_generator_matches_numpy(dim, dr):
    if dim == 2:
        tank, fluid, vert = ((0, 0, 0), (4, 3, 0)), ((0, 0, 0), (1, 2, 0)), 1
    else:
        tank, fluid, vert = ((0, 0, 0), (4, 1, 3)), ((0, 0, 0), (1, 1, 2)), 2
    xf = states.lattice_box(dim, dr, fluid[0], fluid[1])
    assert np.array_equal(xf, np_covering_box(dim, dr, fluid[0], fluid[1]))
    xw = states.lattice_walls(dim, dr, tank[0], tank[1], 3.5 * dr, vert, tank[1][vert])
    assert np.array_equal(xw, np_walls(dim, dr, tank[0], tank[1], 3.5 * dr, vert, tank[1][vert]))


def test_dambreak_2d_counts():
    st = states.dam_break(2, 0.02)
    assert st.n_fluid == 51 * 101
    assert st.n_wall == 1363
    # FP64 rounding of x + dx leaves the right wall with two layers
    xw = st.x[st.n_fluid:]
    assert sorted(set(np.round(xw[xw[:, 0] > 4.0, 0], 9))) == [4.02, 4.04]
    jit = st.x[: st.n_fluid, :2] - np_covering_box(2, 0.02, (0, 0, 0), (1, 2, 0))[:, :2]
    assert np.abs(jit).max() <= 1e-6 * 0.02 and np.abs(jit).max() > 0


def header_functions():
    names = []
    for h in ("sph.h", "sph_batch.h", "sph_setup.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(sph_[a-z_0-9]+)\s*\(", txt, re.M)
    return sorted(set(names))


def test_abi_exports_every_declared_symbol():
    names = header_functions()
    assert "sph_create" in names and "sph_lattice_box" in names
    lib = ctypes.CDLL(N.LIB_PATH)  # loading needs no GPU
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES)


def test_params_struct_layout():
    assert ctypes.sizeof(N.sph_params) == 12 * 8
    p = sp.make_params(m=2.0, g=(0, -9.8, 0), dt=1e-4)
    assert p.m == 2.0 and p.g[1] == -9.8 and p.coeff == 1.0
    with pytest.raises(TypeError):
        sp.make_params(bogus=1.0)


def test_system_desc_layout():
    # double h; double bbox_min[3], bbox_max[3]; int periodic, pad; sph_params
    assert ctypes.sizeof(N.sph_system_desc) == 8 + 48 + 8 + 12 * 8
    assert N.sph_system_desc.params.offset == 64


def test_operator_registry():
    assert sp.OPERATORS["find_rho!"] is sp.find_rho and sp.find_rho.binary
    # names the reference defines with different bodies need their program
    for name in ("find_pressure!", "internal_force!", "move!", "accelerate!"):
        with pytest.raises(KeyError):
            sp._resolve(name)
    cd = sp.program_operators("collapse_dry")
    tc = sp.program_operators("test_collision_2d")
    assert sp._resolve("collapse_dry:move!") is cd["move!"]
    assert cd["move!"].id == N.SPH_OP_MOVE and cd["move!"].fixed == {"coeff": 0.5}
    assert tc["move!"].id == N.SPH_OP_DRIFT and tc["move!"].fixed == {"coeff": 1.0}
    assert cd["accelerate!"].id == N.SPH_OP_ACCELERATE and tc["accelerate!"].id == N.SPH_OP_KICK
    assert cd["find_pressure!"].id == N.SPH_OP_FIND_PRESSURE_CONTINUITY
    assert tc["find_pressure!"].id == N.SPH_OP_FIND_PRESSURE_RHO0
    assert cd["internal_force!"].id == N.SPH_OP_INTERNAL_FORCE
    assert tc["internal_force!"].id == N.SPH_OP_PRESSURE_FORCE_RHO0
    # fixed parameters override the call's
    assert sp._params(cd["move!"], dict(dt=1e-3, coeff=1.0)).coeff == 0.5
    # a Python callable goes to the host fallback
    assert sp._resolve(lambda p, q, r: None) is None
    assert sp._arity(lambda p, q, r: None) == 3 and sp._arity(lambda p: None) == 1
    with pytest.raises(KeyError):
        sp._resolve("no_such_operator!")


def test_host_kernels_match_reference_constants():
    h = 0.3
    assert sp.wendland2(h, 0.0) == pytest.approx(7 / math.pi / h ** 2, rel=1e-15)
    assert sp.wendland3(h, 2 * h) == 0.0
    assert sp.rDwendland3(h, 0.0) == pytest.approx(-210 / math.pi / h ** 5, rel=1e-15)


def test_header_constants_match_python():
    """Every numeric #define of the C headers that the Python mirror also
    names has the same value there."""
    seen = 0
    for h in ("sph.h", "sph_batch.h", "sph_setup.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        for name, val in re.findall(r"^#define\s+(SPH_[A-Z0-9_]+)\s+(-?\d+)\b", txt, re.M):
            if hasattr(N, name):
                assert getattr(N, name) == int(val), name
                seen += 1
    assert seen >= 10


def test_lattice_spec_layout():
    # int dim, kind; double dr, lo[3], hi[3], width; int lid_axis, hydro_axis;
    # double lid, type, rho, P, rho0, g, H, c, jitter; uint64_t seed
    assert ctypes.sizeof(N.sph_lattice_spec) == 160
    assert N.sph_lattice_spec.lid.offset == 80 and N.sph_lattice_spec.seed.offset == 152


def test_jitter_draws_are_counter_based():
    """The device generator draws jitter value t of the splitmix64 stream as
    mix(seed + (t + 1) * gamma); the host's sequential sph_jitter must give
    the same values (the stream is counter-based)."""
    M = (1 << 64) - 1
    gamma = 0x9E3779B97F4A7C15

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    for dim in (2, 3):
        n, amp, seed = 50, 1e-6 * 0.02, 7
        x = np.zeros((n, 3))
        N.lib().sph_jitter(x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, dim, amp, seed)
        ref = np.zeros((n, 3))
        for i in range(n):
            for d in range(dim):
                t = i * dim + d
                u = float(mix((seed + (t + 1) * gamma) & M) >> 11) * 2.0 ** -53
                ref[i, d] = 0.0 + (2.0 * u - 1.0) * amp
        assert np.array_equal(x, ref)


def test_run_count_division_is_exact():
    """The pair kernels' run split counts a target's runs as ceil(pop / B)
    with one FMA, trunc(pop * r + c), r = 1/B and c = (B - 1/2)/B rounded to
    FP32 (a compile-time table in csrc/pair_kernels.h, run_div / runs_of).
    Exact for every pair count pop <= 64 and run length B <= 71, also with r
    one ulp off either way (the margin is 1/(2B), the error below 3e-5/B).
    The FMA's single rounding is emulated by the exact double sum (under 40
    significant bits here) rounded to FP32."""
    pops = np.arange(65, dtype=np.float64)
    for B in range(1, 72):
        r = np.float32(1.0) / np.float32(B)
        c = (np.float32(B) - np.float32(0.5)) / np.float32(B)
        for rr in (r, np.nextafter(r, np.float32(0)), np.nextafter(r, np.float32(1))):
            v = (pops * np.float64(rr) + np.float64(c)).astype(np.float32)
            got = np.trunc(v).astype(np.int64)
            want = (np.arange(65) + B - 1) // B
            assert np.array_equal(got, want), (B, rr)
