"""Host side of the request loop: the seeded request stream and the sizes of
the workload states built for a request (no GPU)."""
import numpy as np
import pytest

from smoothedparticles_hip import serving


def test_request_stream_is_seeded():
    a = serving.make_requests(40, rate=3.0, seed=1)
    b = serving.make_requests(40, rate=3.0, seed=1)
    key = lambda rs: [(r.kind, r.n_target, r.steps, r.arrival, r.re) for r in rs]  # noqa: E731
    assert key(a) == key(b)
    assert key(a) != key(serving.make_requests(40, rate=3.0, seed=2))
    assert all(1000 <= r.n_target <= 500000 and 200 <= r.steps <= 2000 for r in a)
    assert all(50.0 <= r.re <= 1000.0 for r in a)
    assert np.all(np.diff([r.arrival for r in a]) > 0)
    assert {r.kind for r in a} == set(serving.KINDS)


@pytest.mark.parametrize("kind", serving.KINDS)
@pytest.mark.parametrize("n", [1000, 20000])
def test_state_size_follows_request(kind, n):
    st = serving.make_state(kind, n)
    # the fluid (all of a Taylor-Green box) is within 25% of the request
    assert 0.75 * n <= st.n_fluid <= 1.25 * n
    r = serving.Request(0, kind, n, 100, 0.0)
    assert r.footprint() == serving.BYTES_PER_PARTICLE[r.dim] * len(st.x)
    assert r.footprint(1500.5) == int(np.ceil(1500.5 * len(st.x)))
