"""The continuous-batching request loop (not in the reference): every
request completes with its own step count, admission keeps the estimated
device footprint within the budget, and each result has the same bits as the
request run alone."""
import numpy as np
import pytest

import smoothedparticles_hip as sp
from smoothedparticles_hip import serving

pytestmark = pytest.mark.gpu


def alone(r, precision, tiles):
    st = serving.make_state(r.kind, r.n_target, r.re)
    b = sp.Batch(precision)
    b.set_tile(tiles[r.dim])
    sid = b.add_state(st, periodic=r.periodic)
    b.init_rho0()
    done = 0
    while done < r.steps:  # any split of the steps gives the same bits
        k = min(37, r.steps - done)
        b.step(k)
        done += k
    out = serving.summary(st, b, sid)
    b.close()
    return out


@pytest.mark.parametrize("precision", ["mixed"])
def test_served_results_equal_alone(precision):
    print("serving stream", flush=True)
    reqs = serving.make_requests(12, rate=0.0, seed=5, n_lo=800, n_hi=6000, steps_lo=10,
                                 steps_hi=60)
    kinds = {r.kind for r in reqs}
    assert len(kinds) >= 2
    budget = 3 * max(r.footprint() for r in reqs)
    srv = serving.Server(precision, budget=budget, quantum=16)
    fin = srv.run(reqs, realtime=False)
    assert len(fin) == len(reqs)
    assert srv.peak_in_use <= budget
    assert srv.peak_active >= 2  # requests overlapped in a batch
    for r in fin:
        assert r.done_steps == r.steps
        assert r.finished >= r.admitted >= r.arrival
    for r in sorted(fin, key=lambda r: r.rid)[:4]:
        ref = alone(r, precision, srv.tiles)
        assert r.result["n"] == ref["n"]
        assert r.result["kinetic"] == ref["kinetic"], r.rid
        assert r.result["momentum"] == ref["momentum"], r.rid
    st = serving.latency_stats(fin)
    assert st["p99_s"] >= st["p50_s"] > 0
    srv.close()
