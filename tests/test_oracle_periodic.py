"""Periodic systems of the oracle (not in the reference), pinned without the
oracle's own cell code:

* neighbour sets against a numpy minimum-image brute force over all 3^d
  images, on random points with seam cases, boxes with exactly 3 cells per
  axis and positions outside the box (wrapped at the build);
* the density of a perfect periodic lattice against a lattice sum of the
  Wendland kernel written out here;
* Taylor-Green vortex physics: kinetic-energy decay against the closed form
  KE(0) exp(-16 pi^2 nu t) and conservation of momentum.
"""
import math

import numpy as np
import pytest

import oracle as O
from smoothedparticles_hip import states


def min_image_pairs(x, h, lo, hi, axes):
    """All ordered pairs p != q with an image of q within h of p; the image
    coordinate is x_q + k box (k in -1, 0, 1), the distance
    sqrt((d1^2 + d2^2) + d3^2) in FP64.  Asserts that at most one image is
    within h."""
    box = np.asarray(hi, dtype=np.float64) - np.asarray(lo, dtype=np.float64)
    n = len(x)
    shifts = [np.zeros(3)]
    for d in range(axes):
        new = []
        for s in shifts:
            for k in (-1.0, 1.0):
                t = s.copy()
                t[d] = k * box[d]
                new.append(t)
        shifts += new
    hits = np.zeros((n, n), dtype=np.int64)
    for s in shifts:
        xq = x + s  # x_q + shift, one IEEE add per component
        d = x[:, None, :] - xq[None, :, :]
        r = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
        hits += r <= h
    np.fill_diagonal(hits, 0)
    assert hits.max() <= 1
    p, q = np.nonzero(hits)
    return set(zip(p.tolist(), q.tolist()))


def oracle_pairs(s):
    return set(map(tuple, s.pairs().tolist()))


@pytest.mark.parametrize("dim,box,h,n", [(2, (1.0, 1.0), 0.1, 600), (2, (0.9, 0.6), 0.19, 300),
                                         (2, (0.3, 0.45), 0.099, 200), (3, (1.0, 0.8, 0.6), 0.19, 500),
                                         (3, (0.3, 0.3, 0.3), 0.099, 300)])
def test_periodic_pairs_match_minimum_image(dim, box, h, n):
    rng = np.random.Generator(np.random.PCG64(7))
    lo = np.array([-0.2, 0.1, -0.05])
    hi = lo.copy()
    for d in range(dim):
        hi[d] = lo[d] + box[d]
    if dim == 2:
        hi[2] = lo[2]
    x = np.zeros((n, 3))
    x[:, 0] = rng.uniform(lo[0], hi[0], n)
    x[:, 1] = rng.uniform(lo[1], hi[1], n)
    if dim == 3:
        x[:, 2] = rng.uniform(lo[2], hi[2], n)
    else:
        x[:, 2] = lo[2]
    # seam cases: exactly on bmin, just below bmax, outside the box
    x[0, 0] = lo[0]
    x[1, 0] = np.nextafter(hi[0], -np.inf)
    x[2, 1] = hi[1] + 0.3 * box[1]
    x[3, 0] = lo[0] - 1.7 * box[0]
    s = O.OracleSystem(h, lo, hi, periodic=True)
    s.load(x)
    assert s.create_cell_list() == n
    xw = s.field("x").copy()
    # wrapped into [bmin, bmax] on every periodic axis, up to the rounding
    # of x - box floor((x - bmin) / box)
    for d in range(dim):
        assert xw[:, d].min() >= lo[d] - 1e-15 and xw[:, d].max() <= hi[d] + 1e-15
    np.testing.assert_allclose(xw[3, 0], x[3, 0] + 2 * box[0], rtol=0, atol=1e-15)
    ref = min_image_pairs(xw, h, lo, hi, dim)
    got = oracle_pairs(s)
    assert got == ref
    assert all((q, p) in got for p, q in got)


def test_periodic_requires_three_cells():
    with pytest.raises(ValueError):
        O.OracleSystem(0.2, [0, 0, 0], [0.5, 1.0, 0.0], periodic=True)


def test_periodic_removes_non_finite_only():
    s = O.OracleSystem(0.1, [0, 0, 0], [1.0, 1.0, 0.0], periodic=True)
    x = np.array([[0.5, 0.5, 0.0], [np.nan, 0.2, 0.0], [3.25, -0.5, 0.0], [0.1, np.inf, 0.0]])
    s.load(x)
    assert s.create_cell_list() == 2
    np.testing.assert_array_equal(s.field("x"), [[0.5, 0.5, 0.0], [0.25, 0.5, 0.0]])


def wendland2(h, r):
    q = r / h
    return 7.0 / (math.pi * h * h) * (1 - q) ** 4 * (1 + 4 * q) if q <= 1 else 0.0


def test_periodic_lattice_density_is_uniform_lattice_sum():
    n = 24
    st = states.taylor_green(n, 100.0)
    s = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi, periodic=True)
    s.load(st.x)
    s.create_cell_list()
    s.reset_rho()
    s.density_sum(self_term=True, **st.params())
    rho = s.field("rho")
    # lattice sum over all offsets (i, j) dr with |.| <= R, self term included
    k = int(math.ceil(st.R / st.dr)) + 1
    tot = 0.0
    for i in range(-k, k + 1):
        for j in range(-k, k + 1):
            r = math.hypot(i * st.dr, j * st.dr)
            if r <= st.R:
                tot += st.m * wendland2(st.R, r)
    np.testing.assert_allclose(rho, tot, rtol=1e-12)


@pytest.mark.parametrize("n,Re", [(32, 50.0), (32, 200.0)])
def test_taylor_green_decay_and_momentum(n, Re):
    st = states.taylor_green(n, Re)
    pr = st.params()
    s = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi, periodic=True)
    s.load(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    s.create_cell_list()
    s.field("rho0")[:] = 0.0
    s.density_sum_rho0(self_term=True, **pr)
    ke0 = states.taylor_green_kinetic(st, s.field("v"))
    nsteps = 200
    s.step_wcsph(nsteps, use_rho0=True, **pr)
    ke = states.taylor_green_kinetic(st, s.field("v"))
    t = nsteps * st.dt
    rate = -math.log(ke / ke0) / (16 * math.pi ** 2 * st.nu * t)
    assert abs(rate - 1.0) < 0.05, rate
    mom = s.field("v").sum(axis=0) * st.m
    assert np.abs(mom).max() < 1e-14
