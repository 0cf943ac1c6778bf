"""Pins of the oracle's assemble_matrix visits (src/core.jl:196-225): the
cell loop without the p == q exclusion.  CPU only."""
import numpy as np
import pytest

import oracle as O


def rows(a):
    a = np.asarray(a, dtype=np.int64).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


@pytest.mark.parametrize("dim,n,h", [(2, 800, 0.07), (3, 900, 0.2)])
def test_matrix_entries_brute_force(dim, n, h):
    """On a grid that does not alias, the visits are every (i, j) with
    RN(sqrt((d1*d1 + d2*d2) + d3*d3)) <= h, i == j included, once each."""
    rng = np.random.default_rng(n)
    x = rng.random((n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    o = O.OracleSystem(h, np.zeros(3), np.array([1.0, 1.0, 1.0 if dim == 3 else 0.0]))
    o.load(x)
    o.create_cell_list()
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    ref = np.argwhere(r <= h)
    assert np.array_equal(rows(o.matrix_entries()), rows(ref))


def test_matrix_entries_on_an_aliased_grid():
    """A grid one cell wide aliases (di + dj = 0 for three key_diff
    entries): the visits without the diagonal are the pair sweep's visits,
    with the same multiplicities, and each particle meets itself once per
    key_diff entry equal to 0."""
    rng = np.random.default_rng(3)
    x = rng.random((200, 3)) * [0.05, 1.0, 0.0]
    o = O.OracleSystem(0.1, np.zeros(3), np.array([0.05, 1.0, 0.0]))
    o.load(x)
    o.create_cell_list()
    e = o.matrix_entries()
    off = e[e[:, 0] != e[:, 1]]
    assert np.array_equal(rows(off), rows(o.pairs()))
    zeros = sum(1 for d in o.key_geometry()[3] if d == 0)
    assert zeros == 3
    diag = e[e[:, 0] == e[:, 1], 0]
    assert np.array_equal(np.bincount(diag, minlength=200), np.full(200, zeros))
