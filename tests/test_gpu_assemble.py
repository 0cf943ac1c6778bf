"""assemble_matrix / assemble_vector (src/core.jl:175-225) on the device's
neighbour pairs against the oracle's cell-loop visits, p == q included."""
import numpy as np
import pytest
import scipy.sparse

import oracle as O
import smoothedparticles_hip as sp

pytestmark = pytest.mark.gpu


def func(h):
    def f(fields, I, J, r):
        # shaped like examples/collapse_dry_implicit.jl:154-163: a diagonal
        # term of particle i and a kernel-derivative term off the diagonal
        diag = h * h * (1.0 + fields["rho"][I] * 1e-3)
        return np.where(I == J, diag, -2.0 * h * h * (1.0 - r / h) ** 3)
    return f


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dim,n,h,width", [(2, 2500, 0.05, 1.0), (3, 3000, 0.15, 1.0),
                                           (2, 300, 0.1, 0.05)])
def test_assemble_matrix_matches_oracle(precision, dim, n, h, width):
    rng = np.random.default_rng(n)
    x = rng.random((n, 3)) * [width, 1.0, 1.0]
    if dim == 2:
        x[:, 2] = 0.0
    hi = np.array([width, 1.0, 1.0 if dim == 3 else 0.0])
    rho = 1000.0 + rng.random(n)
    g = sp.ParticleSystem(sp.Box(0, 0, 0, *hi), h, precision)
    g.add_particles(x, rho=rho)
    o = O.OracleSystem(h, np.zeros(3), hi)
    o.load(x, None, rho)
    assert sp.create_cell_list(g) == o.create_cell_list() == n
    A = sp.assemble_matrix(g, func(h))
    e = o.matrix_entries()
    d = x[e[:, 0]] - x[e[:, 1]]
    r = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    V = func(h)({"rho": rho}, e[:, 0], e[:, 1], r)
    B = scipy.sparse.csr_matrix((V, (e[:, 0], e[:, 1])), shape=(n, n))
    assert A.nnz == B.nnz
    diff = abs(A - B)
    # distinct entries are single visits and equal exactly; entries visited
    # several times (the aliased grid) may be summed in another order
    tol = 0.0 if width == 1.0 else 1e-15 * abs(B).max()
    assert diff.max() <= tol
    b = sp.assemble_vector(g, lambda f: -f["rho"] / 1e-3)
    assert np.array_equal(b, -rho / 1e-3)
