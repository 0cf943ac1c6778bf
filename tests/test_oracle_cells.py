"""Pins of the oracle's key geometry, cell list and neighbour sets
(src/structs.jl:63-106, src/core.jl:51-112) against closed forms and an
O(N^2) brute force.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def oracle_system(h, lo, hi, x, **fields):
    o = O.OracleSystem(h, np.asarray(lo, float), np.asarray(hi, float))
    o.load(x, **fields)
    return o


@pytest.mark.parametrize("case", GOLDEN["key_geometry"], ids=lambda c: f"h{c['h']}")
def test_key_geometry_golden(case):
    o = O.OracleSystem(case["h"], np.array(case["bmin"]), np.array(case["bmax"]))
    phase, lim, km, diff = o.key_geometry()
    assert list(phase) == case["key_phase"]
    assert list(lim) == case["key_lim"]
    assert km == case["key_max"]
    assert diff == case["key_diff"]


def test_find_key_closed_form_and_nan():
    h = 0.096
    lo, hi = np.array([-0.07, -0.07, 0.0]), np.array([4.07, 3.07, 0.0])
    o = O.OracleSystem(h, lo, hi)
    phase, lim, _, _ = o.key_geometry()
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = np.array([rng.uniform(lo[0], hi[0]), rng.uniform(lo[1], hi[1]), 0.0])
        i = 1 + math.floor(x[0] / h) - phase[0]
        j = 1 + math.floor(x[1] / h) - phase[1]
        assert o.find_key(x) == i + lim[0] * (j - 1)
    assert o.find_key(np.array([np.nan, 0.0, 0.0])) == -1
    assert o.find_key(np.array([0.0, np.inf, 0.0])) == -1


def as_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


@pytest.mark.parametrize("dim,n,h", [(2, 800, 0.08), (3, 1500, 0.15), (2, 3000, 0.031)])
def test_pairs_match_brute_force(dim, n, h):
    rng = np.random.default_rng(dim * 1000 + n)
    x = rng.random((n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    lo, hi = np.zeros(3), np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0])
    o = oracle_system(h, lo, hi, x)
    assert o.create_cell_list() == n
    _, lim, _, _ = o.key_geometry()
    assert lim[0] >= 4 and lim[1] >= 4
    got = o.pairs()
    ref = O.brute_pairs(x, h)
    assert len(got) == len(ref)
    assert as_set(got) == as_set(ref)
    # symmetry of the relation
    assert as_set(got) == as_set(got[:, ::-1])


@pytest.mark.parametrize("dim,k", [(2, 2.4), (2, 4.8), (3, 2.4), (3, 4.8)])
def test_lattice_interior_counts(dim, k):
    expected = next(c["count"] for c in GOLDEN["lattice_counts"]
                    if c["dim"] == dim and c["R_over_dr"] == k)
    dr = 0.01
    m = 21 if dim == 2 else 13
    idx = np.arange(m) * dr
    if dim == 2:
        g = np.stack(np.meshgrid(idx, idx, indexing="ij"), -1).reshape(-1, 2)
        x = np.concatenate([g, np.zeros((len(g), 1))], 1)
    else:
        x = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3)
    h = k * dr
    hi = np.array([idx[-1]] * 3)
    if dim == 2:
        hi[2] = 0.0
    o = oracle_system(h, np.zeros(3), hi, x)
    o.create_cell_list()
    counts, _ = o.neighbour_counts()
    centre = np.argmin(((x - idx[m // 2]) ** 2)[:, :dim].sum(1))
    assert counts[centre] == expected


def test_exact_tie_and_duplicates():
    h = 0.5
    # r == h exactly is accepted (src/core.jl:105 rejects only r > h)
    x = np.array([[1.0, 1.0, 0.0], [1.5, 1.0, 0.0],
                  [3.0, 3.0, 0.0], [np.nextafter(3.5, 4.0), 3.0, 0.0],
                  [2.0, 0.25, 0.0], [2.0, 0.25, 0.0]])  # duplicate positions
    o = oracle_system(h, np.zeros(3), np.array([4.0, 4.0, 0.0]), x)
    o.create_cell_list()
    s = as_set(o.pairs())
    assert (0, 1) in s and (1, 0) in s
    assert (2, 3) not in s and (3, 2) not in s
    assert (4, 5) in s and (5, 4) in s          # distinct particles at r = 0
    assert all(p != q for p, q in s)            # p is never its own neighbour
    assert s == as_set(O.brute_pairs(x, h))


def test_small_grid_aliasing_multiplicity():
    """With key_lim[1] = 2 the 9 key offsets alias, so the reference visits
    some cells twice and applies the operator twice (src/core.jl:96-102)."""
    h = 0.25
    lo, hi = np.zeros(3), np.array([0.4, 1.0, 0.0])
    rng = np.random.default_rng(5)
    x = np.zeros((60, 3))
    x[:, 0] = rng.uniform(0, 0.4, 60)
    x[:, 1] = rng.uniform(0, 1.0, 60)
    o = oracle_system(h, lo, hi, x)
    o.create_cell_list()
    phase, lim, km, diff = o.key_geometry()
    assert lim[0] == 2
    keys = [1 + math.floor(p[0] / h) - phase[0] + lim[0] * (math.floor(p[1] / h) - phase[1])
            for p in x]
    expected = {}
    for p in range(len(x)):
        for q in range(len(x)):
            if p == q or math.sqrt((x[p, 0] - x[q, 0]) ** 2 + (x[p, 1] - x[q, 1]) ** 2) > h:
                continue
            mult = sum(1 for d in diff if 1 <= keys[p] + d <= km and keys[p] + d == keys[q])
            if mult:
                expected[(p, q)] = mult
    got = {}
    for p, q in o.pairs().tolist():
        got[(p, q)] = got.get((p, q), 0) + 1
    assert got == expected
    assert max(got.values()) >= 2


def test_removal_swap_with_end():
    """Particles outside the bounding box (and NaN) are removed by
    particles[r_i] = particles[end+1-i] over descending r_i
    (src/core.jl:63-81); restated here on a Python list of ids."""
    rng = np.random.default_rng(7)
    n = 40
    x = rng.random((n, 3))
    x[:, 2] = 0.0
    bad = [3, 17, 38, 39, 21]
    x[3, 0] = 1.5
    x[17, 1] = np.nan
    x[38, 0] = -0.1
    x[39, 1] = 2.0
    x[21, 0] = np.inf
    o = oracle_system(0.2, np.zeros(3), np.array([1.0, 1.0, 0.0]), x)
    assert o.create_cell_list() == n - len(bad)
    ids = list(range(n))
    rem = sorted(bad, reverse=True)
    for i, r in enumerate(rem, start=1):
        ids[r] = ids[n - i]
    ids = ids[: n - len(rem)]
    assert list(o.field("id")) == ids
    xs = x[ids]
    assert as_set(o.pairs()) == as_set(O.brute_pairs(xs, 0.2))
