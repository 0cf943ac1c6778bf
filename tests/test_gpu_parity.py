"""Parity of the HIP library (through its C ABI) with the FP64 oracle.

Neighbour sets are compared bit-exactly (as multisets of reference-index
pairs) in both precision modes.  Tolerances, stated per test:
  FP64 pair operators  rel 1e-12 (sum order and the @fastmath kernel
                       lowering differ at the ulp level)
  mixed pair operators rel 5e-6 for rho, 1e-4 of max|a| for forces
  unary operators      bit-exact (same FP64 expressions)
"""
import numpy as np
import pytest

import oracle as O
import smoothedparticles_hip as sp
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu

PHYS = dict(m=0.4, rho0=1000.0, c=50.0, alpha=0.02, hs=0.05, mu=0.0, nu=0.0,
            g=(0.0, -9.8, 0.0), dt=1e-4)


def make_both(x, h, lo, hi, precision="fp64", v=None, rho=None, P=None, type=None):
    g = sp.ParticleSystem(sp.Box(*lo, *hi), h, precision)
    g.add_particles(x, v=v, rho=rho, P=P, type=type)
    o = O.OracleSystem(h, np.asarray(lo, float), np.asarray(hi, float))
    o.load(x, v, rho, P, type)
    return g, o


def sorted_pairs(p):
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def random_cloud(n, dim, seed, extent=1.0):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)) * extent
    if dim == 2:
        x[:, 2] = 0.0
    lo = np.zeros(3)
    hi = np.array([extent, extent, extent if dim == 3 else 0.0])
    return x, lo, hi, rng


def fields(rng, n, dim):
    v = rng.normal(size=(n, 3))
    if dim == 2:
        v[:, 2] = 0.0
    rho = 1000.0 + 20.0 * rng.random(n)
    P = 2500.0 * (rho - 1000.0) + 100.0 * rng.normal(size=n)
    typ = (rng.random(n) < 0.2).astype(float)
    return v, rho, P, typ


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim,n,h", [(2, 3000, 0.05), (3, 4000, 0.12), (2, 2000, 0.021)])
def test_pairs_random(precision, dim, n, h):
    x, lo, hi, _ = random_cloud(n, dim, 11 + n)
    g, o = make_both(x, h, lo, hi, precision)
    assert sp.create_cell_list(g) == o.create_cell_list() == n
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim", [2, 3])
def test_pairs_far_from_origin(precision, dim):
    """A box far from the origin (coordinates ~1e3, ulp ~1e-13) holding a
    lattice of spacing h/2, so many pairs sit at exactly h or within an ulp
    of it: the tile cull and the mixed guard band must keep every pair the
    FP64 predicate accepts."""
    h = 0.1
    rng = np.random.default_rng(5 + dim)
    k = 14 if dim == 3 else 40
    axes = [np.arange(k) * (h / 2)] * dim
    g0 = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim)
    jit = rng.integers(-2, 3, size=g0.shape) * np.spacing(1000.0)
    x = np.zeros((len(g0), 3))
    x[:, :dim] = 1000.0 + g0 + jit
    lo = np.full(3, 999.9)
    hi = np.array([1000.0 + k * h / 2 + 0.1] * 3)
    if dim == 2:
        lo[2] = hi[2] = 0.0
    g, o = make_both(x, h, lo, hi, precision)
    assert sp.create_cell_list(g) == o.create_cell_list() == len(x)
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_ties_duplicates_boundaries(precision):
    h = 0.25
    pts = [[1.0, 1.0, 0], [1.25, 1.0, 0], [2.0, 2.0, 0], [np.nextafter(2.25, 3), 2.0, 0],
           [0.5, 0.5, 0], [0.5, 0.5, 0], [0.5, 0.75, 0], [0.75, 0.5, 0], [0.0, 0.0, 0],
           [2.5, 2.5, 0], [2.5, 2.75, 0], [np.nextafter(0.25, 0), 0.0, 0]]
    rng = np.random.default_rng(2)
    # lattice points exactly on cell faces at distances sqrt(k) * h / 2
    for i in range(9):
        for j in range(9):
            pts.append([i * 0.125 + 0.5, j * 0.125 + 0.5, 0.0])
    x = np.array(pts, float)
    g, o = make_both(x, h, np.zeros(3), np.array([3.0, 3.0, 0.0]), precision)
    sp.create_cell_list(g)
    o.create_cell_list()
    got = sorted_pairs(sp.neighbour_pairs(g))
    assert np.array_equal(got, sorted_pairs(o.pairs()))
    s = set(map(tuple, got.tolist()))
    assert (0, 1) in s and (2, 3) not in s and (4, 5) in s


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lim_case", ["2d_L1_2", "3d_L2_2"])
def test_pairs_small_grid_aliasing(precision, lim_case):
    """Aliased key offsets visit some pairs twice, as the reference does."""
    rng = np.random.default_rng(9)
    if lim_case == "2d_L1_2":
        h, lo, hi = 0.25, np.zeros(3), np.array([0.4, 1.0, 0.0])
        x = np.zeros((80, 3))
        x[:, 0] = rng.uniform(0, 0.4, 80)
        x[:, 1] = rng.uniform(0, 1.0, 80)
    else:
        h, lo, hi = 0.3, np.array([-1.0, 0.0, -1.0]), np.array([1.0, 0.5, 1.0])
        x = rng.uniform(lo, hi, (150, 3))
    g, o = make_both(x, h, lo, hi, precision)
    sp.create_cell_list(g)
    o.create_cell_list()
    ref = sorted_pairs(o.pairs())
    assert len(ref) > len(set(map(tuple, ref.tolist())))  # multiplicities exist
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairs_2d_grid_with_varying_z(precision):
    """key_lim[3] == 1 but z varies inside the box: 9 offsets, 3D distance."""
    rng = np.random.default_rng(4)
    x = rng.random((2000, 3))
    x[:, 2] *= 0.01
    lo, hi = np.zeros(3), np.array([1.0, 1.0, 0.01])
    g, o = make_both(x, 0.05, lo, hi, precision)
    assert g.key_geometry()[1][2] == 1
    sp.create_cell_list(g)
    o.create_cell_list()
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("R_over_dr", [2.4, 4.8])
def test_pairs_dambreak_2d(precision, R_over_dr):
    st = states.dam_break(2, 0.02, R_over_dr=R_over_dr)
    g, o = make_both(st.x, st.R, st.bbox_lo, st.bbox_hi, precision, st.v, st.rho, st.P, st.type)
    sp.create_cell_list(g)
    o.create_cell_list()
    assert g.key_geometry() == o.key_geometry()
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 5e-6)])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("self_term", [True, False])
def test_density_sum(precision, tol, dim, self_term):
    n = 3000 if dim == 2 else 5000
    x, lo, hi, rng = random_cloud(n, dim, 21 + dim)
    h = 0.06 if dim == 2 else 0.15
    v, rho, P, typ = fields(rng, n, dim)
    g, o = make_both(x, h, lo, hi, precision, v, rho, P, typ)
    sp.create_cell_list(g)
    o.create_cell_list()
    sp.apply(g, sp.reset_rho)
    sp.apply(g, sp.find_rho, self=self_term, m=0.37)
    o.reset_rho()
    o.density_sum(self_term, m=0.37)
    ref = o.field("rho")
    assert np.max(np.abs(g.get("rho") - ref) / np.abs(ref)) <= tol
    # rho0 variant accumulates onto the existing field
    sp.apply(g, sp.find_rho0, self=self_term, m=0.37)
    o.density_sum_rho0(self_term, m=0.37)
    ref0 = o.field("rho0")
    assert np.max(np.abs(g.get("rho0") - ref0) / np.abs(ref0)) <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-4)])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mu", [0.0, 1e-3])
def test_internal_force(precision, tol, dim, mu):
    n = 3000 if dim == 2 else 5000
    x, lo, hi, rng = random_cloud(n, dim, 31 + dim)
    h = 0.06 if dim == 2 else 0.15
    v, rho, P, typ = fields(rng, n, dim)
    g, o = make_both(x, h, lo, hi, precision, v, rho, P, typ)
    sp.create_cell_list(g)
    o.create_cell_list()
    kw = dict(PHYS, mu=mu, hs=h / 2)
    sp.apply(g, sp.reset_a)
    sp.apply(g, sp.internal_force, **kw)
    o.reset_a()
    o.internal_force(**kw)
    ref = o.field("a")
    got = g.get("a")
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.max(np.abs(got - ref)) <= tol * scale
    assert np.all(got[typ != 0.0] == 0.0)  # walls receive no force


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", 1e-5)])
@pytest.mark.parametrize("dim", [2, 3])
def test_pressure_force_rho0_and_continuity(precision, tol, dim):
    """internal_force! with rho0^2 (tests/test_collision_2d.jl:75-78) and
    balance_of_mass! (examples/collapse_dry.jl:112-115); the mixed forms are
    FP32 pair arithmetic with FP64 accumulation: within 1e-5 of max|a| and
    max|Drho|, FP64 within 1e-11."""
    n = 2500 if dim == 2 else 4000
    x, lo, hi, rng = random_cloud(n, dim, 41)
    v, rho, P, typ = fields(rng, n, dim)
    g, o = make_both(x, 0.06 if dim == 2 else 0.12, lo, hi, precision, v, rho, P, typ)
    sp.create_cell_list(g)
    o.create_cell_list()
    kw = dict(PHYS, nu=1e-3)
    sp.apply(g, sp.internal_force_rho0, **kw)
    o.pressure_force_rho0(**kw)
    ref = o.field("a")
    assert np.max(np.abs(g.get("a") - ref)) <= tol * np.abs(ref).max()
    sp.apply(g, sp.balance_of_mass, **kw)
    o.balance_of_mass(**kw)
    ref = o.field("Drho")
    assert np.max(np.abs(g.get("Drho") - ref)) <= tol * np.abs(ref).max()


def test_unary_operators_bit_exact():
    n = 5000
    x, lo, hi, rng = random_cloud(n, 3, 51)
    v, rho, P, typ = fields(rng, n, 3)
    g, o = make_both(x, 0.1, lo, hi, "fp64", v, rho, P, typ)
    a = rng.normal(size=(n, 3))
    g.set("a", a)
    o.field("a")[:] = a
    d = rng.normal(size=n)
    g.set("Drho", d)
    o.field("Drho")[:] = d
    kw = dict(PHYS, coeff=0.5, g=(0.1, -9.8, 0.3))
    sp.apply(g, sp.accelerate, **kw); o.accelerate(**kw)
    # test_collision_2d.jl's move! and accelerate!: every particle, a kept
    sp.apply(g, sp.drift, **kw); o.drift(**kw)
    sp.apply(g, sp.kick, **kw); o.kick(**kw)
    for f in ("x", "v", "a"):
        assert np.array_equal(g.get(f), o.field(f)), f
    assert np.array_equal(g.get("a"), a)
    # collapse_dry.jl's move!: p.Dv = VEC0 for every particle, fluid drift
    sp.apply(g, sp.move, **kw); o.move(**kw)
    assert np.all(g.get("a") == 0.0) and np.all(o.field("a") == 0.0)
    sp.apply(g, sp.find_pressure, **kw); o.eos(**kw)
    for f in ("x", "v", "P"):
        assert np.array_equal(g.get(f), o.field(f)), f
    sp.apply(g, sp.find_pressure_continuity, **kw); o.eos_continuity(**kw)
    for f in ("rho", "Drho", "P"):
        assert np.array_equal(g.get(f), o.field(f)), f
    r0 = 1000 + rng.random(n)
    g.set("rho0", r0); o.field("rho0")[:] = r0
    sp.apply(g, sp.find_pressure_rho0, **kw); o.eos(use_field=True, **kw)
    assert np.array_equal(g.get("P"), o.field("P"))
    sp.apply(g, sp.reset_rho); sp.apply(g, sp.reset_a)
    assert np.all(g.get("rho") == 0) and np.all(g.get("a") == 0)


def test_reductions():
    n = 20000
    x, lo, hi, rng = random_cloud(n, 3, 61)
    v, rho, P, typ = fields(rng, n, 3)
    g, o = make_both(x, 0.1, lo, hi, "fp64", v, rho, P, typ)
    kw = dict(PHYS, g=(0, 0, -9.8))
    assert sp.energy(g, **kw) == pytest.approx(o.energy(**kw), rel=1e-12)
    r0 = 1000 + rng.random(n)
    g.set("rho0", r0); o.field("rho0")[:] = r0
    assert sp.energy(g, "collision", **kw) == pytest.approx(o.energy("collision", **kw), rel=1e-12)
    assert np.allclose(sp.momentum(g, **kw), o.momentum(**kw), rtol=1e-12, atol=1e-12)


def test_removal_matches_reference_order():
    n = 3000
    x, lo, hi, rng = random_cloud(n, 2, 71)
    v, rho, P, typ = fields(rng, n, 2)
    bad = rng.choice(n, 37, replace=False)
    x[bad[:10], 0] = 1.5
    x[bad[10:20], 1] = np.nan
    x[bad[20:30], 0] = -np.inf
    x[bad[30:], 1] = -0.01
    x[n - 1, 0] = 7.0            # the last particle too
    g, o = make_both(x, 0.05, lo, hi, "fp64", v, rho, P, typ)
    assert sp.create_cell_list(g) == o.create_cell_list() == len(g)
    for f in ("x", "v", "rho", "P", "type"):
        assert np.array_equal(g.get(f), o.field(f), equal_nan=True), f
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))
    # a second build without removals keeps the order
    assert sp.create_cell_list(g) == o.create_cell_list()
    assert np.array_equal(g.get("x"), o.field("x"))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dense_cell_and_large_removal_bucket(precision):
    """Buckets larger than the rank's LDS staging: a cell holding 3000
    particles and a removal bucket of 2500, beside ordinary cells (the blocks
    that straddle them rank in global memory).  Fields come back in the
    reference's order, the neighbour digests and the density equal the
    oracle's."""
    rng = np.random.default_rng(97)
    spread = rng.random((2000, 3))
    dense = 0.5 + 0.05 * rng.random((3000, 3))
    out = rng.random((2500, 3))
    out[:, 0] = 2.0                       # outside the box: removed
    x = np.concatenate([spread, dense, out])
    x[:, 2] = 0.0
    x = x[rng.permutation(len(x))]
    lo, hi = np.zeros(3), np.array([1.0, 1.0, 0.0])
    v, rho, P, typ = fields(rng, len(x), 2)
    g, o = make_both(x, 0.05, lo, hi, precision, v, rho, P, typ)
    assert sp.create_cell_list(g) == o.create_cell_list() == 5000
    for f in ("x", "v", "rho", "type"):
        assert np.array_equal(g.get(f), o.field(f)), f
    cg, hg = sp.neighbour_digest(g)
    co, ho = o.neighbour_digest()
    assert np.array_equal(cg, co) and np.array_equal(hg, ho)
    sp.apply(g, sp.reset_rho)
    sp.apply(g, sp.find_rho, self=True, m=0.37)
    o.reset_rho()
    o.density_sum(True, m=0.37)
    tol = 1e-12 if precision == "fp64" else 5e-6
    assert np.max(np.abs(g.get("rho") / o.field("rho") - 1)) <= tol


def test_stale_cell_list_is_an_error():
    x, lo, hi, _ = random_cloud(500, 2, 81)
    g, _ = make_both(x, 0.1, lo, hi)
    sp.create_cell_list(g)
    sp.apply(g, sp.move, dt=1e-3)
    with pytest.raises(sp.SPHError):
        sp.apply(g, sp.find_rho, m=1.0)


@pytest.mark.parametrize("T", [8, 16, 32, 64])
def test_tile_sizes_agree(T):
    st = states.dam_break(2, 0.02)
    g, o = make_both(st.x, st.R, st.bbox_lo, st.bbox_hi, "fp64", st.v, st.rho, st.P, st.type)
    g.set_tile(T)
    sp.create_cell_list(g)
    o.create_cell_list()
    assert np.array_equal(sorted_pairs(sp.neighbour_pairs(g)), sorted_pairs(o.pairs()))
    sp.apply(g, sp.reset_rho)
    sp.apply(g, sp.find_rho, self=True, m=st.m)
    o.reset_rho()
    o.density_sum(True, m=st.m)
    assert np.max(np.abs(g.get("rho") / o.field("rho") - 1)) <= 1e-13


def test_autotune_keeps_state():
    st = states.dam_break(2, 0.02)
    g, _ = make_both(st.x, st.R, st.bbox_lo, st.bbox_hi, "fp64", st.v, st.rho, st.P, st.type)
    sp.create_cell_list(g)
    before = {f: g.get(f) for f in ("rho", "a", "x")}
    T = g.autotune(**st.params())
    assert T in (8, 16, 32, 64)
    for f, a in before.items():
        assert np.array_equal(g.get(f), a)


def test_operators_after_mixed_density_pass():
    """After a mixed density pass the rho0^2 force and balance_of_mass! (in
    their mixed forms) find the same pairs whether they read its masks (with
    or without the kept-candidate lists) or re-test, and add them in the same
    runs: equal to 1e-12 of the largest value."""
    n = 2500
    x, lo, hi, rng = random_cloud(n, 2, 57)
    v, rho, P, typ = fields(rng, n, 2)
    kw = dict(PHYS, nu=1e-3)
    out = []
    for masks, kept in ((False, True), (True, True), (True, False)):
        g, _ = make_both(x, 0.06, lo, hi, "mixed", v, rho, P, typ)
        g.set_pair_masks(masks)
        g.set_kept_lists(kept)
        sp.create_cell_list(g)
        sp.apply(g, sp.reset_rho)
        sp.apply(g, sp.find_rho, self=True, m=0.37)
        sp.apply(g, sp.reset_a)
        sp.apply(g, sp.internal_force_rho0, **kw)
        sp.apply(g, sp.balance_of_mass, **kw)
        out.append((g.get("a"), g.get("Drho")))
    for o in out[1:]:
        for k in range(2):
            assert np.max(np.abs(o[k] - out[0][k])) <= 1e-12 * np.abs(out[0][k]).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim", [2, 3])
def test_pair_masks_match_retesting(precision, dim):
    """Forces read from the density pass's accepted-candidate masks equal
    the forces of a pass that re-tests every candidate, bit for bit, with
    and without the kept-candidate lists."""
    n = 3000 if dim == 2 else 6000
    x, lo, hi, rng = random_cloud(n, dim, 91 + dim)
    h = 0.06 if dim == 2 else 0.15
    v, rho, P, typ = fields(rng, n, dim)
    out = []
    for masks, kept in ((True, True), (False, True), (True, False)):
        g, _ = make_both(x, h, lo, hi, precision, v, rho, P, typ)
        g.set_pair_masks(masks)
        g.set_kept_lists(kept)
        sp.create_cell_list(g)
        sp.apply(g, sp.reset_rho)
        sp.apply(g, sp.find_rho, self=True, m=0.37)
        sp.apply(g, sp.find_pressure, c=50.0, rho0=1000.0)
        sp.apply(g, sp.reset_a)
        sp.apply(g, sp.internal_force, **dict(PHYS, hs=h / 2))
        out.append((g.get("rho"), g.get("a")))
    for o in out[1:]:
        assert np.array_equal(out[0][0], o[0])
        assert np.array_equal(out[0][1], o[1])
