"""Pins of the oracle's kernels and operators against closed forms
(src/kernels.jl, tests/test_kernels.jl, examples/collapse_dry.jl,
tests/test_collision_2d.jl).  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
L = O.lib()
KC = GOLDEN["kernel_consts"]


def simpson(f, a, b, n=4000):
    h = (b - a) / n
    s = f(a) + f(b)
    for i in range(1, n):
        s += (4 if i % 2 else 2) * f(a + i * h)
    return s * h / 3


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_invariants(dim):
    """tests/test_kernels.jl:20-43, with a finer quadrature."""
    W = L.orc_wendland2 if dim == 2 else L.orc_wendland3
    D = L.orc_Dwendland2 if dim == 2 else L.orc_Dwendland3
    rD = L.orc_rDwendland2 if dim == 2 else L.orc_rDwendland3
    h = 0.42
    assert W(h, 4.0) == 0.0 and D(h, 4.0) == 0.0 and rD(h, 4.0) == 0.0
    assert math.isfinite(W(h, 0.0)) and math.isfinite(rD(h, 0.0))
    if dim == 2:
        integral = simpson(lambda r: 2 * math.pi * r * W(h, r), 0.0, h)
    else:
        integral = simpson(lambda r: 4 * math.pi * r * r * W(h, r), 0.0, h)
    assert integral == pytest.approx(1.0, rel=1e-9)
    assert simpson(lambda r: D(h, r), 0.2, 0.3) == pytest.approx(W(h, 0.3) - W(h, 0.2), rel=1e-9)
    assert rD(h, 0.1) == pytest.approx(D(h, 0.1) / 0.1, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_kernel_closed_form(dim):
    h = 0.37
    x = 0.5
    cw = KC["w2"] if dim == 2 else KC["w3"]
    cd = KC["dw2"] if dim == 2 else KC["dw3"]
    W = L.orc_wendland2 if dim == 2 else L.orc_wendland3
    rD = L.orc_rDwendland2 if dim == 2 else L.orc_rDwendland3
    assert W(h, x * h) == pytest.approx(cw * 0.5 ** 4 * 3.0 / h ** dim, rel=1e-14)
    assert W(h, 0.0) == pytest.approx(cw / h ** dim, rel=1e-14)
    assert rD(h, x * h) == pytest.approx(-cd * 0.5 ** 3 / h ** (dim + 2), rel=1e-14)
    assert W(h, h) == 0.0


def lattice(dim, m, dr, origin=0.0):
    idx = origin + np.arange(m) * dr
    grids = np.meshgrid(*([idx] * dim), indexing="ij")
    x = np.stack(grids, -1).reshape(-1, dim)
    if dim == 2:
        x = np.concatenate([x, np.zeros((len(x), 1))], 1)
    return x


@pytest.mark.parametrize("dim,k", [(2, 4.8), (3, 2.4)])
def test_density_lattice_sum(dim, k):
    """rho of an interior lattice particle = m * sum of W over the lattice
    offsets within R (a finite sum, evaluated here with math.fsum)."""
    dr = 0.02
    R = k * dr
    m_ = 1000.0 * dr ** dim
    n = 25 if dim == 2 else 13
    x = lattice(dim, n, dr)
    hi = x.max(0)
    o = O.OracleSystem(R, np.zeros(3), hi)
    o.load(x)
    o.create_cell_list()
    o.density_sum(True, m=m_)
    centre = np.argmin(((x - (n // 2) * dr) ** 2)[:, :dim].sum(1))
    kk = int(math.ceil(k))
    terms = []
    C = KC["w2"] if dim == 2 else KC["w3"]
    rng3 = range(-kk, kk + 1) if dim == 3 else (0,)
    for i in range(-kk, kk + 1):
        for j in range(-kk, kk + 1):
            for l in rng3:
                r = math.sqrt(i * i + j * j + l * l) * dr
                if r <= R:
                    q = r / R
                    terms.append(m_ * C * (1 - q) ** 4 * (1 + 4 * q) / R ** dim)
    assert o.field("rho")[centre] == pytest.approx(math.fsum(terms), rel=1e-12)


def test_isolated_particle_self_density():
    o = O.OracleSystem(0.1, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o.load(np.array([[0.5, 0.5, 0.0], [0.1, 0.1, 0.0]]))
    o.create_cell_list()
    o.density_sum(True, m=2.0)
    assert np.allclose(o.field("rho"), 2.0 * KC["w2"] / 0.01, rtol=1e-14)


def two_particles(dx, v=(0.0, 0.0), P=(1000.0, 500.0), rho=(1001.0, 999.0), types=(0.0, 0.0)):
    x = np.array([[1.0, 1.0, 0.0], [1.0 + dx, 1.0, 0.0]])
    vel = np.array([[v[0], 0.0, 0.0], [v[1], 0.0, 0.0]])
    o = O.OracleSystem(0.1, np.zeros(3), np.array([2.0, 2.0, 0.0]))
    o.load(x, vel, np.array(rho), np.array(P), np.array(types))
    o.create_cell_list()
    return o


def test_pressure_force_two_particles():
    m, h, dx = 0.4, 0.1, 0.06
    o = two_particles(dx)
    o.internal_force(m=m, rho0=1000.0, c=50.0)
    u = 1 - dx / h
    rdw = -KC["dw2"] * u ** 3 / h ** 4
    A = 1000.0 / 1001.0 ** 2 + 500.0 / 999.0 ** 2
    a0 = -m * rdw * A * (-dx)
    assert o.field("a")[0, 0] == pytest.approx(a0, rel=1e-13)
    assert o.field("a")[1, 0] == pytest.approx(-m * rdw * A * dx, rel=1e-13)
    # walls receive no force (examples/collapse_dry.jl:136)
    o = two_particles(dx, types=(0.0, 1.0))
    o.internal_force(m=m, rho0=1000.0, c=50.0)
    assert o.field("a")[1, 0] == 0.0 and o.field("a")[0, 0] != 0.0


def test_artificial_viscosity_two_particles():
    m, h, dx, c, alpha, hs = 0.4, 0.1, 0.06, 50.0, 0.02, 0.05
    P = (0.0, 0.0)
    # approaching: v_pq . x_pq < 0 -> Monaghan term
    o = two_particles(dx, v=(1.0, -1.0), P=P)
    o.internal_force(m=m, rho0=1000.0, c=c, alpha=alpha, hs=hs)
    u = 1 - dx / h
    rdw = -KC["dw2"] * u ** 3 / h ** 4
    vx = (1.0 - -1.0) * (-dx)
    rhobar = 0.5 * (1001.0 + 999.0)
    Pi = -alpha * c * hs * vx / (rhobar * (dx * dx + 0.01 * hs * hs))
    assert o.field("a")[0, 0] == pytest.approx(-m * Pi * rdw * (-dx), rel=1e-12)
    # receding: no viscosity, and P = 0 -> zero force
    o = two_particles(dx, v=(-1.0, 1.0), P=P)
    o.internal_force(m=m, rho0=1000.0, c=c, alpha=alpha, hs=hs)
    assert np.all(o.field("a") == 0.0)


def test_momentum_antisymmetry():
    rng = np.random.default_rng(3)
    n = 600
    x = rng.random((n, 3))
    x[:, 2] = 0.0
    o = O.OracleSystem(0.08, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o.load(x, rng.normal(size=(n, 3)) * [1, 1, 0], 1000 + rng.random(n), rng.random(n) * 1e4,
           np.zeros(n))
    o.create_cell_list()
    o.internal_force(m=0.01, rho0=1000.0, c=50.0, alpha=0.1, hs=0.04)
    a = o.field("a")
    tot = a.sum(0)
    assert np.abs(tot).max() <= 1e-12 * np.abs(a).sum()


def test_continuity_two_particles():
    m, h, dx, nu = 0.4, 0.1, 0.06, 1e-3
    o = two_particles(dx, v=(1.0, -0.5))
    o.balance_of_mass(m=m, nu=nu)
    u = 1 - dx / h
    rdw = -KC["dw2"] * u ** 3 / h ** 4
    d = (-dx) * (1.0 + 0.5)
    assert o.field("Drho")[0] == pytest.approx(m * rdw * (d + 2 * nu * (1001.0 - 999.0)), rel=1e-13)


def test_free_fall_closed_form():
    """Isolated particles under gravity with symplectic Euler:
    v_n = v0 + n dt g, x_n = x0 + n dt v0 + dt^2 g n(n+1)/2."""
    x0 = np.array([[0.5, 2.0, 0.0], [3.0, 1.0, 0.0]])
    v0 = np.array([[0.3, 0.0, 0.0], [-0.2, 0.5, 0.0]])
    o = O.OracleSystem(0.01, np.zeros(3), np.array([4.0, 3.0, 0.0]))
    o.load(x0, v0, np.full(2, 1000.0), np.zeros(2), np.zeros(2))
    g = np.array([0.0, -9.8, 0.0])
    dt, n = 1e-3, 50
    for _ in range(n):
        o.create_cell_list()
        o.reset_a()
        o.accelerate(dt=dt, g=g, coeff=1.0)
        o.move(dt=dt, coeff=1.0)
    assert np.allclose(o.field("v"), v0 + n * dt * g, rtol=0, atol=1e-13)
    xn = x0 + n * dt * v0 + dt * dt * g * n * (n + 1) / 2
    assert np.allclose(o.field("x"), xn, rtol=0, atol=1e-13)


def test_energy_closed_form():
    x = np.array([[0.1, 0.2, 0.0], [1.0, 2.0, 0.0]])
    v = np.array([[1.0, 2.0, 0.0], [0.0, -1.0, 0.0]])
    rho = np.array([1010.0, 990.0])
    o = O.OracleSystem(0.05, np.zeros(3), np.array([4.0, 3.0, 0.0]))
    o.load(x, v, rho, np.zeros(2), np.zeros(2))
    m, c, rho0, g = 0.4, 50.0, 1000.0, np.array([0.0, -9.8, 0.0])
    e = 0.0
    for i in range(2):
        e += 0.5 * m * v[i] @ v[i] - m * g @ x[i] + m * c * c * (
            math.log(rho[i] / rho0) + rho0 / rho[i] - 1.0)
    assert o.energy(m=m, c=c, rho0=rho0, g=g) == pytest.approx(e, rel=1e-14)
    o.field("rho0")[:] = [1000.5, 999.0]
    ec = sum(0.5 * m * v[i] @ v[i] + 0.5 * m * c * c * (rho[i] - o.field("rho0")[i]) ** 2 / rho0 ** 2
             for i in range(2))
    assert o.energy("collision", m=m, c=c, rho0=rho0) == pytest.approx(ec, rel=1e-14)


def test_eos():
    o = O.OracleSystem(0.05, np.zeros(3), np.array([1.0, 1.0, 0.0]))
    o.load(np.zeros((3, 3)), rho=np.array([990.0, 1000.0, 1010.0]))
    o.eos(c=50.0, rho0=1000.0)
    assert list(o.field("P")) == [2500.0 * -10.0, 0.0, 2500.0 * 10.0]
