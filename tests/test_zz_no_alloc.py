"""sph_probe_sum_noalloc (include/sph.h): the probe sums of sph_probe_sum
staged in the library's static device array instead of a per-call
allocation.  Called through the C ABI and compared with sph_probe_sum bit
for bit, twice in a row (the second call smaller, so stale contents of the
array would show), in both precision modes; an oversized call is rejected
on the host."""
import ctypes

import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import _native as N
from smoothedparticles_hip import states

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def call(fn, g, params, flags, x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    out = np.full((x.shape[0], 4), -7.0)
    rc = fn(g._h, ctypes.byref(params), flags, x.ctypes.data_as(_dp), x.shape[0],
            out.ctypes.data_as(_dp))
    return rc, out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_probe_sum_equals_allocating_call(precision):
    st = states.dam_break(3, 0.04)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=np.random.default_rng(1).normal(size=st.x.shape), rho=st.rho,
                    P=st.P, type=st.type)
    sp.create_cell_list(g)
    L = N.lib()
    params = sp.make_params(**p)
    rng = np.random.default_rng(2)
    span = st.bbox_hi - st.bbox_lo
    for np_, flags in ((50000, N.SPH_PROBE_FLUID_ONLY), (3000, 0)):
        x = st.bbox_lo + rng.random((np_, 3)) * span
        x[:5] = st.x[:5]                 # at particles
        x[5] = [np.nan, 0.5, 0.5]        # key -1 as in the reference
        rc1, ref = call(L.sph_probe_sum, g, params, flags, x)
        rc2, got = call(L.sph_probe_sum_noalloc, g, params, flags, x)
        assert rc1 == 0 and rc2 == 0, N.lib().sph_last_error()
        assert np.array_equal(got, ref, equal_nan=True)
        assert np.count_nonzero(ref[:, 0]) > 0
    # above the stated maximum: a clear error, nothing launched
    big = np.zeros((N.SPH_PROBE_NOALLOC_MAX + 1, 3))
    rc, _ = call(L.sph_probe_sum_noalloc, g, params, 0, big)
    assert rc == N.SPH_EINVAL
    assert b"SPH_PROBE_NOALLOC_MAX" in L.sph_last_error()
    # the handle still works after the rejected call
    rc, got = call(L.sph_probe_sum_noalloc, g, params, 0, st.x[:10])
    rc1, ref = call(L.sph_probe_sum, g, params, 0, st.x[:10])
    assert rc == 0 and rc1 == 0 and np.array_equal(got, ref)
