"""The continuous-batching request loop (not in the reference): every
request completes with its own step count, admission keeps the device
footprint within the budget, each result has the same bits as the request
run alone, small served requests match the FP64 oracle, and a full-size
stream (48 requests of 1k-500k particles, 200-2000 steps) runs through."""
import numpy as np
import pytest

import oracle as O
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
    assert all(srv.measured[d] for d in {r.dim for r in reqs})
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


def test_served_requests_match_the_oracle():
    """Six small requests of every kind served together in the mixed mode;
    each final state against the FP64 oracle run of the same request:
    rho rel <= 1e-5, velocity within 3e-4 of max |v|, same particle count.
    (A dam break 20-40 steps after rest has small velocities, set by the
    near balance of pressure and gravity, so the FP32 force error shows
    most there: ~1e-4 of max |v|, falling to 6e-5 by step 100.)"""
    print("serving vs oracle", flush=True)
    reqs = serving.make_requests(6, rate=0.0, seed=11, n_lo=1000, n_hi=3000, steps_lo=20,
                                 steps_hi=40)
    assert {r.kind for r in reqs} == set(serving.KINDS)
    srv = serving.Server("mixed", budget=1 << 30, quantum=8, keep=("rho", "v"))
    fin = srv.run(reqs, realtime=False)
    assert srv.peak_active >= 3
    for r in fin:
        st = serving.make_state(r.kind, r.n_target, r.re)
        p = st.params()
        o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi, periodic=r.periodic)
        o.load(st.x, st.v, st.rho, st.P, st.type)
        o.create_cell_list()
        o.density_sum_rho0(True, **p)
        o.step_wcsph(r.steps, use_rho0=True, **p)
        assert r.result["n"] == o.n
        er = np.max(np.abs(r.result["rho"] / o.field("rho") - 1))
        vo = o.field("v")
        ev = np.max(np.abs(r.result["v"] - vo)) / max(np.abs(vo).max(), 1e-5)
        print(f"  {r.kind} n={o.n} steps={r.steps}: rho {er:.1e} v {ev:.1e}")
        assert er <= 1e-5 and ev <= 3e-4
    srv.close()


def test_full_size_stream_properties():
    """The benchmark's stream shape: 48 requests, N log-uniform in
    [1k, 500k], 200-2000 steps, all kinds, served back to back in the mixed
    mode under a 24 GB budget."""
    print("serving full-size stream", flush=True)
    reqs = serving.make_requests(48, rate=0.0, seed=0)
    budget = 24 << 30
    srv = serving.Server("mixed", budget=budget, quantum=50)
    fin = srv.run(reqs, realtime=False)
    assert len(fin) == 48
    assert srv.peak_in_use <= budget
    assert srv.peak_active >= 4
    for r in fin:
        assert r.done_steps == r.steps
        assert r.result["n"] > 0 and np.isfinite(r.result["kinetic"])
        if r.kind == "taylor_green":
            assert np.all(np.abs(r.result["momentum"]) <= 1e-9 * max(1.0, r.result["n"]))
    st = serving.latency_stats(fin)
    print(f"full stream: {srv.particle_steps:.3e} particle-steps in {srv.wall:.1f} s, "
          f"{srv.pairs / srv.wall:.3e} pair-int/s, peak {srv.peak_active} requests, "
          f"peak footprint {srv.peak_footprint / 1e9:.2f} GB, p50 {st['p50_s']:.2f} s")
    srv.close()
