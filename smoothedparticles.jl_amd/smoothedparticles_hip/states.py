"""Initial states of the benchmark workloads, generated on the host with the
lattice routines of libsph_hip (include/sph_setup.h).

Dam break (2D and 3D), after examples/collapse_dry.jl:88-102 on a square /
cubic lattice: a fluid block in a tank, wall particles from
BoundaryLayer(tank, grid, 3.5 dr) without the lid, hydrostatic P and rho
for every particle (collapse_dry.jl:97-100), and a seeded uniform jitter
of +-1e-6 dr on the fluid positions.

Taylor-Green vortex (not in the reference): a periodic unit box, square
lattice, u = sin(2 pi x) cos(2 pi y), v = -cos(2 pi x) sin(2 pi y).
"""
import ctypes
import math

import numpy as np

from . import _native as N
from ._native import lib

_dptr = ctypes.POINTER(ctypes.c_double)


def _arr3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


def lattice_box(dim, dr, lo, hi):
    lo, hi = _arr3(lo), _arr3(hi)
    L = lib()
    n = L.sph_lattice_box(dim, dr, lo.ctypes.data_as(_dptr), hi.ctypes.data_as(_dptr), None, 0)
    if n < 0:
        raise ValueError("invalid lattice arguments")
    out = np.empty((max(n, 1), 3), dtype=np.float64)
    L.sph_lattice_box(dim, dr, lo.ctypes.data_as(_dptr), hi.ctypes.data_as(_dptr),
                      out.ctypes.data_as(_dptr), n)
    return out[:n]


def lattice_walls(dim, dr, tank_lo, tank_hi, width, lid_axis, lid):
    lo, hi = _arr3(tank_lo), _arr3(tank_hi)
    L = lib()
    args = (dim, dr, lo.ctypes.data_as(_dptr), hi.ctypes.data_as(_dptr), width, lid_axis, lid)
    n = L.sph_lattice_wall_layer(*args, None, 0)
    if n < 0:
        raise ValueError("invalid lattice arguments")
    out = np.empty((max(n, 1), 3), dtype=np.float64)
    L.sph_lattice_wall_layer(*args, out.ctypes.data_as(_dptr), n)
    return out[:n]


def jitter(x, dim, amplitude, seed):
    x = np.ascontiguousarray(x)
    lib().sph_jitter(x.ctypes.data_as(_dptr), x.shape[0], dim, amplitude, seed)
    return x


def lattice_spec(kind, dim, dr, lo, hi, width=0.0, lid_axis=-1, lid=0.0, type=0.0, rho=0.0,
                 P=0.0, hydro=None, jitter=0.0, seed=0):
    """An sph_lattice_spec (include/sph_setup.h) for the device generator.
    hydro = (axis, rho0, g, H, c) gives the hydrostatic state of
    examples/collapse_dry.jl:97-100 instead of the constant rho and P."""
    sp_ = N.sph_lattice_spec()
    sp_.dim, sp_.kind, sp_.dr = dim, kind, float(dr)
    for d in range(3):
        sp_.lo[d], sp_.hi[d] = float(lo[d]), float(hi[d])
    sp_.width, sp_.lid_axis, sp_.lid = float(width), int(lid_axis), float(lid)
    sp_.type, sp_.rho, sp_.P = float(type), float(rho), float(P)
    if hydro is None:
        sp_.hydro_axis = -1
    else:
        sp_.hydro_axis = int(hydro[0])
        sp_.rho0, sp_.g, sp_.H, sp_.c = (float(v) for v in hydro[1:])
    sp_.jitter, sp_.seed = float(jitter), int(seed)
    return sp_


class State:
    """Particle arrays (reference order), domain box and physical params."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def params(self):
        return dict(m=self.m, rho0=self.rho0, c=self.c, alpha=self.alpha, hs=self.hs,
                    mu=0.0, nu=0.0, g=tuple(self.g), dt=self.dt, coeff=1.0)


def dam_break_specs(dim, dr, R_over_dr=4.8, c=50.0, rho0=1000.0, alpha=0.02, gravity=9.8,
                    jitter_rel=1e-6, seed=0):
    """The dam break of dam_break() for the device generator: a State with
    the geometry and parameters but no particle arrays, and the lattice
    specs of the fluid and of the walls (generate them in that order)."""
    if dim == 2:
        tank = ((0.0, 0.0, 0.0), (4.0, 3.0, 0.0))
        fluid = ((0.0, 0.0, 0.0), (1.0, 2.0, 0.0))
        vert = 1
    else:
        tank = ((0.0, 0.0, 0.0), (4.0, 1.0, 3.0))
        fluid = ((0.0, 0.0, 0.0), (1.0, 1.0, 2.0))
        vert = 2
    width = 3.5 * dr
    H = fluid[1][vert]
    hydro = (vert, rho0, -gravity, H, c)
    fs = lattice_spec(N.SPH_LATTICE_BOX, dim, dr, fluid[0], fluid[1], type=0.0, hydro=hydro,
                      jitter=jitter_rel * dr, seed=seed)
    ws = lattice_spec(N.SPH_LATTICE_WALL_LAYER, dim, dr, tank[0], tank[1], width=width,
                      lid_axis=vert, lid=tank[1][vert], type=1.0, hydro=hydro)
    lo = np.array(tank[0], dtype=np.float64)
    hi = np.array(tank[1], dtype=np.float64)
    for d in range(dim):
        lo[d] = tank[0][d] - width
        hi[d] = tank[1][d] + width
    g = np.zeros(3)
    g[vert] = -gravity
    hs = 2.4 * dr
    st = State(dim=dim, dr=dr, bbox_lo=lo, bbox_hi=hi, R=R_over_dr * dr, hs=hs,
               m=rho0 * dr ** dim, rho0=rho0, c=c, alpha=alpha, g=g, dt=0.1 * hs / c)
    return st, fs, ws


def dam_break(dim, dr, R_over_dr=4.8, c=50.0, rho0=1000.0, alpha=0.02, gravity=9.8,
              jitter_rel=1e-6, seed=0, fluid=None, tank=None):
    """Dam-break initial state.

    2D: tank [0,4]x[0,3], fluid [0,1]x[0,2], y vertical.
    3D: tank [0,4]x[0,1]x[0,3], fluid [0,1]x[0,1]x[0,2], z vertical.
    Smoothing length hs = 2.4 dr; neighbour radius R = R_over_dr * dr
    (4.8 dr = the support 2 hs of the Wendland kernel); dt = 0.1 hs / c;
    m = rho0 dr^dim.
    """
    if dim == 2:
        tank = tank or ((0.0, 0.0, 0.0), (4.0, 3.0, 0.0))
        fluid = fluid or ((0.0, 0.0, 0.0), (1.0, 2.0, 0.0))
        vert = 1
    else:
        tank = tank or ((0.0, 0.0, 0.0), (4.0, 1.0, 3.0))
        fluid = fluid or ((0.0, 0.0, 0.0), (1.0, 1.0, 2.0))
        vert = 2
    width = 3.5 * dr
    xf = lattice_box(dim, dr, fluid[0], fluid[1])
    xf = jitter(xf, dim, jitter_rel * dr, seed)
    xw = lattice_walls(dim, dr, tank[0], tank[1], width, vert, tank[1][vert])
    x = np.concatenate([xf, xw], axis=0)
    typ = np.concatenate([np.zeros(len(xf)), np.ones(len(xw))])
    g = np.zeros(3)
    g[vert] = -gravity
    H = fluid[1][vert]
    # examples/collapse_dry.jl:97-100, applied to every particle
    P = rho0 * g[vert] * (x[:, vert] - H)
    rho = rho0 + P / (c * c)
    lo = np.array(tank[0], dtype=np.float64)
    hi = np.array(tank[1], dtype=np.float64)
    for d in range(dim):
        lo[d] = tank[0][d] - width
        hi[d] = tank[1][d] + width
    hs = 2.4 * dr
    return State(dim=dim, dr=dr, x=x, v=np.zeros_like(x), rho=rho, P=P, type=typ,
                 n_fluid=len(xf), n_wall=len(xw), bbox_lo=lo, bbox_hi=hi,
                 R=R_over_dr * dr, hs=hs, m=rho0 * dr ** dim, rho0=rho0, c=c, alpha=alpha,
                 g=g, dt=0.1 * hs / c)


def collision_2d(dr=2.0e-2, c=20.0, v0=1.0):
    """Two colliding discs of tests/test_collision_2d.jl:16-57: square
    lattice, radius 0.4, centres (-0.5,-0.1) and (0.5,0.1), velocities +-v0,
    h = 2.4 dr (support), domain [-10,10]^2, dt = 0.1 h / c."""
    R = 2.4 * dr

    def disc(cx, cy, r):
        lo = (cx - r, cy - r, 0.0)
        hi = (cx + r, cy + r, 0.0)
        # covering(Squaregrid, Circle): bbox lattice, then the circle test
        # (x1 - c1)^2 + (x2 - c2)^2 <= r^2 (src/geometry.jl:62-64)
        i0, i1 = math.floor(lo[0] / dr), math.ceil(hi[0] / dr)
        j0, j1 = math.floor(lo[1] / dr), math.ceil(hi[1] / dr)
        pts = []
        for i in range(i0, i1 + 1):
            for j in range(j0, j1 + 1):
                x1, x2 = i * dr, j * dr
                if (x1 - cx) * (x1 - cx) + (x2 - cy) * (x2 - cy) <= r * r:
                    pts.append((x1, x2, 0.0))
        return np.array(pts, dtype=np.float64)

    x1 = disc(-0.5, -0.1, 0.4)
    x2 = disc(0.5, 0.1, 0.4)
    x = np.concatenate([x1, x2])
    v = np.zeros_like(x)
    v[: len(x1), 0] = v0
    v[len(x1):, 0] = -v0
    return State(dim=2, dr=dr, x=x, v=v, rho=np.zeros(len(x)), P=np.zeros(len(x)),
                 type=np.zeros(len(x)), bbox_lo=np.array([-10.0, -10.0, 0.0]),
                 bbox_hi=np.array([10.0, 10.0, 0.0]), R=R, hs=R, m=1000.0 * dr * dr,
                 rho0=1000.0, c=c, alpha=0.0, g=np.zeros(3), dt=0.1 * R / c)


def taylor_green(n, Re, R_over_dr=4.8, c=10.0, rho0=1.0):
    """2D Taylor-Green vortex on an n x n lattice of the periodic unit box.

    dr = 1/n, particles at ((i + 1/2) dr, (j + 1/2) dr); U = L = 1, so the
    kinematic viscosity is nu = 1/Re, applied through the laminar term of the
    internal force (mu = rho0 nu).  c = 10 U (Mach 0.1), hs = 2.4 dr, neighbour
    radius R = R_over_dr dr, dt = 0.1 hs / c, m = rho0 dr^2, no gravity, no
    artificial viscosity.  The kinetic energy of the exact solution decays
    as KE(t) = KE(0) exp(-16 pi^2 nu t).
    """
    dr = 1.0 / n
    g = (np.arange(n, dtype=np.float64) + 0.5) * dr
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = np.zeros((n * n, 3))
    x[:, 0] = X.ravel()
    x[:, 1] = Y.ravel()
    v = np.zeros_like(x)
    tp = 2.0 * math.pi
    v[:, 0] = np.sin(tp * x[:, 0]) * np.cos(tp * x[:, 1])
    v[:, 1] = -np.cos(tp * x[:, 0]) * np.sin(tp * x[:, 1])
    hs = 2.4 * dr
    nu = 1.0 / Re
    st = State(dim=2, dr=dr, x=x, v=v, rho=np.full(n * n, rho0), P=np.zeros(n * n),
               type=np.zeros(n * n), n_fluid=n * n, n_wall=0,
               bbox_lo=np.zeros(3), bbox_hi=np.array([1.0, 1.0, 0.0]), R=R_over_dr * dr,
               hs=hs, m=rho0 * dr * dr, rho0=rho0, c=c, alpha=0.0, g=np.zeros(3),
               dt=0.1 * hs / c, Re=Re, nu=nu, mu=rho0 * nu)
    st.params = lambda: dict(m=st.m, rho0=rho0, c=c, alpha=0.0, hs=hs, mu=st.mu, nu=0.0,
                             g=(0.0, 0.0, 0.0), dt=st.dt, coeff=1.0)
    return st


def taylor_green_kinetic(st, v):
    """Kinetic energy sum m |v|^2 / 2 of a Taylor-Green state."""
    return 0.5 * st.m * float(np.sum(v * v))


def taylor_green_batch(count=256, n_lo=2000, n_hi=60000, re_lo=50.0, re_hi=1000.0, seed=0):
    """The periodic batch workload: particle counts N_i drawn log-uniform in
    [n_lo, n_hi] (lattice side round(sqrt(N_i))) and Re_i uniform in
    [re_lo, re_hi], from a seeded PCG64 stream.  Returns [(side, Re), ...]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), size=count))
    Re = rng.uniform(re_lo, re_hi, size=count)
    return [(int(round(math.sqrt(float(a)))), float(b)) for a, b in zip(N, Re)]
