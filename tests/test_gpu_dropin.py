"""The reference programs' own loops through the drop-in, against the oracle.

- examples/collapse_dry.jl:202-211 driven operator by operator with the
  program's registered operators (its move! resets Dv, its find_pressure! is
  the continuity form), and the fused scheme that does the same on the
  device, FP64 and mixed.
- The host fallback for operators that are not registered: a Python closure
  run over the device's neighbour pairs in the reference's visiting order.
- Neighbour-set digests (count and hash per particle) against the oracle.
"""
import numpy as np
import pytest

import oracle as O
import smoothedparticles_hip as sp
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu


def cd_params(st):
    """collapse_dry.jl's viscosity mu and density diffusion nu (:52-53), no
    artificial viscosity, on the dam-break geometry."""
    p = st.params()
    p.update(mu=8.4e-4, nu=1.0e-6, alpha=0.0)
    return p


def load(st, precision):
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
    o.load(st.x, st.v, st.rho, st.P, st.type)
    return g, o


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def oracle_cd_init(o, p):
    o.create_cell_list()
    o.internal_force(**p)


def test_collapse_dry_loop_through_registered_operators():
    """The reference loop, call for call, with collapse_dry's own operators,
    against the oracle's restatement of the same loop: 100 iterations of the
    2D dam break (dr = 0.02, 6514 particles), FP64."""
    st = states.dam_break(2, 0.02)
    p = cd_params(st)
    g, o = load(st, "fp64")
    ops = sp.program_operators("collapse_dry")
    sp.create_cell_list(g)
    sp.apply(g, ops["internal_force!"], **p)
    oracle_cd_init(o, p)
    for _ in range(100):
        sp.apply(g, ops["accelerate!"], **p)
        sp.apply(g, ops["move!"], **p)
        assert np.all(g.get("a") == 0.0) if _ == 0 else True
        sp.create_cell_list(g)
        sp.apply(g, ops["balance_of_mass!"], **p)
        sp.apply(g, ops["find_pressure!"], **p)
        sp.apply(g, ops["move!"], **p)
        sp.create_cell_list(g)
        sp.apply(g, ops["internal_force!"], **p)
        sp.apply(g, ops["accelerate!"], **p)
    o.step_collapse_dry(100, **p)
    assert len(g) == o.n
    print("collapse_dry 100 its: rho", relmax(g.get("rho"), o.field("rho")),
          "v", relmax(g.get("v"), o.field("v")), "x", relmax(g.get("x"), o.field("x")))
    assert relmax(g.get("rho"), o.field("rho")) <= 1e-12
    assert relmax(g.get("v"), o.field("v")) <= 1e-8
    assert relmax(g.get("x"), o.field("x")) <= 1e-12
    assert sp.energy(g, "wcsph", **p) == pytest.approx(o.energy("wcsph", **p), rel=1e-11)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_collapse_dry_fused_equals_operators(precision):
    """The fused scheme gives the same bits as the operator sequence."""
    st = states.dam_break(2, 0.05)
    p = cd_params(st)
    g1, _ = load(st, precision)
    g2, _ = load(st, precision)
    ops = sp.program_operators("collapse_dry")
    sp.collapse_dry_init(g1, **p)
    sp.collapse_dry_init(g2, **p)
    n1 = sp.step(g1, 5, scheme="collapse_dry", **p)
    n2 = 0
    for _ in range(5):
        sp.apply(g2, ops["accelerate!"], **p)
        sp.apply(g2, ops["move!"], **p)
        sp.create_cell_list(g2)
        sp.apply(g2, ops["balance_of_mass!"], **p)
        n2 += len(sp.neighbour_pairs(g2))
        sp.apply(g2, ops["find_pressure!"], **p)
        sp.apply(g2, ops["move!"], **p)
        sp.create_cell_list(g2)
        sp.apply(g2, ops["internal_force!"], **p)
        sp.apply(g2, ops["accelerate!"], **p)
    assert n1 == n2
    for f in ("x", "v", "a", "rho", "P", "Drho"):
        assert np.array_equal(g1.get(f), g2.get(f)), f


def test_collapse_dry_fp64_fused_vs_oracle_1000_steps():
    """C0 geometry, the reference's own scheme, 1000 fused FP64 steps."""
    st = states.dam_break(2, 0.02)
    p = cd_params(st)
    g, o = load(st, "fp64")
    sp.collapse_dry_init(g, **p)
    oracle_cd_init(o, p)
    E0 = o.energy("wcsph", **p)
    ng = sp.step(g, 1000, scheme="collapse_dry", **p)
    no = o.step_collapse_dry(1000, **p)
    er = relmax(g.get("rho"), o.field("rho"))
    ev = relmax(g.get("v"), o.field("v"))
    ee = abs(sp.energy(g, "wcsph", **p) / o.energy("wcsph", **p) - 1)
    print(f"collapse_dry fp64 1000 steps: pairs {ng} vs {no}; rho {er:.2e} v {ev:.2e} "
          f"E {ee:.2e}; oracle energy change {o.energy('wcsph', **p) / E0 - 1:.2e}")
    assert len(g) == o.n == len(st.x)
    assert abs(ng - no) <= 1e-6 * no
    assert er <= 1e-9 and ee <= 1e-9 and ev <= 1e-5


def test_collapse_dry_mixed_vs_oracle_1000_steps():
    """The mixed scheme (FP32 balance_of_mass! and internal force) against the
    FP64 oracle after 1000 steps: rho rel <= 1e-5, total energy rel <= 1e-4."""
    st = states.dam_break(2, 0.02)
    p = cd_params(st)
    g, o = load(st, "mixed")
    sp.collapse_dry_init(g, **p)
    oracle_cd_init(o, p)
    errs = []
    for k in range(10):
        sp.step(g, 100, scheme="collapse_dry", **p)
        o.step_collapse_dry(100, **p)
        er = relmax(g.get("rho"), o.field("rho"))
        ee = abs(sp.energy(g, "wcsph", **p) / o.energy("wcsph", **p) - 1)
        errs.append((er, ee))
    print("collapse_dry mixed rho/E error per 100 steps:",
          " ".join(f"{a:.1e}/{b:.1e}" for a, b in errs))
    assert len(g) == o.n
    assert max(e[0] for e in errs) <= 1e-5
    assert max(e[1] for e in errs) <= 1e-4


# ---------------------------------------------------------- host fallback --

def find_rho_closure(m, h, dim):
    """tests/test_collision_2d.jl:63-65 written as a Python closure, with the
    Wendland kernel in the oracle's operation order (src/kernels.jl:108-163)."""
    C = 2.228169203286535 if dim == 2 else 3.3422538049298023

    def W(r):
        x = r / h
        if x > 1.0:
            return 0.0
        u = 1.0 - x
        u2 = u * u
        u4 = u2 * u2
        return C * u4 * (1.0 + 4.0 * x) / ((h * h) if dim == 2 else (h * h * h))

    def find_rho(p, q, r):
        p.rho += m * W(r)

    return find_rho


@pytest.mark.parametrize("case", ["dam2d", "cloud3d", "aliased"])
def test_host_fallback_binary_closure_is_the_reference_sum(case):
    """A closure that is not registered runs on the host over the device's
    pairs in the reference's visiting order, so its sum equals the oracle's
    sequential sum bit for bit (self term included)."""
    if case == "dam2d":
        st = states.dam_break(2, 0.05)
        x, h, lo, hi, dim = st.x, st.R, st.bbox_lo, st.bbox_hi, 2
    elif case == "cloud3d":
        rng = np.random.default_rng(5)
        x = rng.random((1500, 3)) * [1.0, 0.8, 0.6]
        h, lo, hi, dim = 0.12, np.zeros(3), np.array([1.0, 0.8, 0.6]), 3
    else:  # key_lim = 2 along x: key_diff entries alias, pairs are visited twice
        rng = np.random.default_rng(6)
        x = rng.random((300, 3)) * [0.35, 1.0, 0.0]
        h, lo, hi, dim = 0.2, np.zeros(3), np.array([0.35, 1.0, 0.0]), 2
    m = 0.37
    g = sp.ParticleSystem(sp.Box(*lo, *hi), h, "fp64")
    g.add_particles(x)
    o = O.OracleSystem(h, lo, hi)
    o.load(x)
    sp.create_cell_list(g)
    o.create_cell_list()
    o.density_sum(True, m=m)
    sp.apply(g, find_rho_closure(m, h, dim), self=True)
    assert np.array_equal(g.get("rho"), o.field("rho"))
    # the visiting order itself: the oracle lists the pairs in it
    assert np.array_equal(sp.visit_order(g), o.pairs())


def test_host_fallback_unary_closure():
    st = states.dam_break(2, 0.05)
    p = st.params()
    g1, _ = load(st, "fp64")
    g2, _ = load(st, "fp64")
    c, rho0 = p["c"], p["rho0"]

    def find_pressure(q):  # examples/collapse_dry.jl:123-127
        q.rho += q.Drho * p["dt"]
        q.Drho = 0.0
        q.P = c ** 2 * (q.rho - rho0)

    g1.set("Drho", np.linspace(-1.0, 1.0, len(g1)))
    g2.set("Drho", np.linspace(-1.0, 1.0, len(g2)))
    sp.apply(g1, find_pressure)
    sp.apply(g2, sp.find_pressure_continuity, **p)
    for f in ("rho", "Drho", "P"):
        assert np.array_equal(g1.get(f), g2.get(f)), f


def test_host_fallback_forwards_parameters_and_vectors():
    """apply! forwards keyword parameters to the closure (src/core.jl:154);
    vector fields are written whole (p.Dv += ...)."""
    st = states.dam_break(2, 0.05)
    g, o = load(st, "fp64")
    sp.create_cell_list(g)
    o.create_cell_list()

    def pull(p, q, r, k=0.0):
        p.Dv = p.Dv + k * (q.x - p.x)

    sp.apply(g, pull, k=2.0)
    a = np.zeros((o.n, 3))
    x = o.field("x")
    for pi, qi in o.pairs():
        a[pi] = a[pi] + 2.0 * (x[qi] - x[pi])
    assert np.array_equal(g.get("a"), a)


# ------------------------------------------------------- neighbour digest --

def digest_equal(g, o):
    cg, hg = sp.neighbour_digest(g)
    co, ho = o.neighbour_digest()
    return np.array_equal(cg, co) and np.array_equal(hg, ho), int(cg.sum())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_digest_random_clouds(precision):
    rng = np.random.default_rng(9)
    for dim, n, h in ((2, 4000, 0.03), (3, 6000, 0.09)):
        x = rng.random((n, 3))
        if dim == 2:
            x[:, 2] = 0.0
        lo, hi = np.zeros(3), np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0])
        g = sp.ParticleSystem(sp.Box(*lo, *hi), h, precision)
        g.add_particles(x)
        o = O.OracleSystem(h, lo, hi)
        o.load(x)
        sp.create_cell_list(g)
        o.create_cell_list()
        ok, npairs = digest_equal(g, o)
        assert ok and npairs == len(o.pairs())


def test_digest_dambreak3d_steps_0_1_10_100():
    """C1's check at a size the oracle runs in seconds: the 3D dam break at
    dr = 0.04 stepped on the device (FP64); at steps 0, 1, 10 and 100 the
    device's neighbour sets equal the oracle's on the same positions, and the
    device state follows the oracle's own run."""
    st = states.dam_break(3, 0.04)
    p = st.params()
    g, o = load(st, "fp64")
    sp.init_reference_density(g, **p)
    o.create_cell_list()
    o.density_sum_rho0(True, **p)
    done = 0
    for k in (0, 1, 10, 100):
        sp.step(g, k - done, scheme="wcsph_se_rho0", **p)
        o.step_wcsph(k - done, use_rho0=True, **p)
        done = k
        sp.create_cell_list(g)
        o2 = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
        o2.load(g.get("x"))
        o2.create_cell_list()
        ok, npairs = digest_equal(g, o2)
        assert ok, f"step {k}"
        print(f"step {k}: {npairs} pairs, rho rel {relmax(g.get('rho'), o.field('rho')):.1e}")
        if k > 0:
            assert relmax(g.get("rho"), o.field("rho")) <= 1e-9
            assert relmax(g.get("x"), o.field("x")) <= 1e-12


def test_digest_dambreak3d_full_c1():
    """C1 at full size (dr = 0.0125, 1.73M particles): per-particle neighbour
    counts and hashes against the oracle at step 0 and after one FP64 step."""
    st = states.dam_break(3, 0.0125)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "fp64")
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    sp.init_reference_density(g, **p)
    for k in (0, 1):
        if k:
            sp.step(g, 1, scheme="wcsph_se_rho0", **p)
            sp.create_cell_list(g)
        o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
        o.load(g.get("x"))
        o.create_cell_list()
        ok, npairs = digest_equal(g, o)
        print(f"C1 step {k}: {len(g)} particles, {npairs} pairs")
        assert ok
