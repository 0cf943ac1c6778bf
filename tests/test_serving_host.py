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


class _FakeBatch:
    """Stands in for a device batch: a live particle count and the bytes
    the library would report for it (capacity kept when particles leave)."""

    def __init__(self):
        self.n = 0
        self.bytes = 0

    def size(self):
        return 1, self.n

    def footprint(self):
        return self.bytes


def test_admission_charge_ignores_capacity_kept_from_retired_requests():
    """The per-particle charge is measured only when the batch reaches a new
    largest particle count: after a large request retires, the batch keeps
    its arrays, and footprint / n would otherwise charge every later request
    as if it needed the retired one's memory."""
    srv = serving.Server.__new__(serving.Server)
    srv.bpp = dict(serving.BYTES_PER_PARTICLE)
    srv.measured = {2: False, 3: False}
    srv.n_high = {2: 0, 3: 0}
    srv.peak_footprint = 0
    b = _FakeBatch()
    srv.batches = {3: b}
    b.n, b.bytes = 400_000, 400_000 * 900      # a large request admitted
    srv._measure(3)
    assert srv.measured[3] and srv.bpp[3] == 900
    b.n = 2_000                                # it retired; its memory stays
    srv._measure(3)
    assert srv.bpp[3] == 900                   # not 180,000 B per particle
    b.n, b.bytes = 500_000, 500_000 * 850      # a new high-water mark
    srv._measure(3)
    assert srv.bpp[3] == 850
    assert srv.peak_footprint == 500_000 * 850
