"""Continuous-batching request loop over Batch handles (not in the reference).

Requests are independent simulations (2D/3D dam break, periodic
Taylor-Green) that arrive over time.  The server keeps one device batch per
dimension; a request is admitted as soon as it has arrived and its
estimated device footprint fits the memory budget, its state stays on the
device while it runs, and it leaves the batch when its step count is done.
Requests that arrive while others run join the running batch between two
stepping calls, so their work shares every launch with the others.

Each batch is advanced in quanta of at most ``quantum`` steps, cut short so
that no request runs past its step count; a request's trajectory therefore
does not depend on what else is in the batch (the batch gives the same
bits as a batch of one with the same tile size).  The 2D and the 3D batch
are stepped from two host threads, each on its own stream, so their
kernels overlap on the device.

Admission is by device footprint: a request is charged its particle count
times the bytes per particle its batch holds, measured with
sph_batch_footprint after an admission that brings the batch to a new
largest particle count (a starting value until the first measurement).  A
batch keeps its device arrays when requests retire, so the ratio is taken
only at that high-water mark, where it measures what the particles need
rather than capacity left over from larger past batches.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Batch, states

KINDS = ("dam_break_2d", "dam_break_3d", "taylor_green")

# Starting value of the admission charge per particle, before a batch of the
# dimension has been measured: particle columns, two buffers, build scratch
# and pair masks, with the capacity doubling of the particle arrays.
BYTES_PER_PARTICLE = {2: 1200, 3: 2400}


class Request:
    """One simulation request."""

    def __init__(self, rid, kind, n_target, steps, arrival, re=None):
        self.rid = rid
        self.kind = kind
        self.n_target = n_target
        self.steps = steps
        self.arrival = arrival
        self.re = re
        self.state = None
        self.sid = None
        self.done_steps = 0
        self.admitted = None
        self.finished = None
        self.result = None

    @property
    def dim(self):
        return 3 if self.kind == "dam_break_3d" else 2

    @property
    def periodic(self):
        return self.kind == "taylor_green"

    def build(self):
        """The initial state (built on the host when first needed)."""
        if self.state is None:
            self.state = make_state(self.kind, self.n_target, self.re)
        return self.state

    @property
    def n(self):
        return len(self.build().x)

    def footprint(self, bytes_per_particle=None):
        bpp = BYTES_PER_PARTICLE[self.dim] if bytes_per_particle is None else bytes_per_particle
        return int(math.ceil(bpp * self.n))

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def make_state(kind, n_target, re=None):
    """A workload state of about n_target particles."""
    if kind == "taylor_green":
        side = max(16, int(round(math.sqrt(n_target))))
        return states.taylor_green(side, re if re is not None else 100.0)
    if kind == "dam_break_2d":
        # fluid lattice (1/dr + 1)(2/dr + 1) ~ 2/dr^2; the walls come on top
        return states.dam_break(2, math.sqrt(2.0 / n_target))
    if kind == "dam_break_3d":
        return states.dam_break(3, (2.0 / n_target) ** (1.0 / 3.0))
    raise ValueError(kind)


def make_requests(count, rate, seed=0, n_lo=1000, n_hi=500000, steps_lo=200, steps_hi=2000,
                  kinds=KINDS):
    """A seeded request stream: Poisson arrivals of the given rate (per
    second), particle counts log-uniform in [n_lo, n_hi], step counts uniform
    in [steps_lo, steps_hi], kinds uniform, Taylor-Green Re uniform in
    [50, 1000]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        kind = kinds[int(rng.integers(len(kinds)))]
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        re = float(rng.uniform(50.0, 1000.0))
        out.append(Request(i, kind, n, steps, t, re))
    return out


def summary(st, b, sid, fields=()):
    """Result of a finished request: kinetic energy and momentum, plus the
    final arrays of `fields`."""
    v = b.get(sid, "v")
    out = {"n": int(len(v)), "kinetic": 0.5 * st.m * float(np.sum(v * v)),
           "momentum": (st.m * v.sum(axis=0)).tolist()}
    for f in fields:
        out[f] = b.get(sid, f)
    return out


class Server:
    """Continuous batching of requests on one GPU.

    precision   "mixed" or "fp64"
    budget      device bytes the admitted requests may use (estimate)
    quantum     steps per stepping call of a batch
    tiles       targets per tile for each dimension (fixed, so a request's
                results do not depend on the batch it ran in)
    keep        fields whose final arrays each result keeps
    graph       replay each stepping call's steps from a hipGraph
    """

    def __init__(self, precision="mixed", budget=16 << 30, quantum=50, tiles=None, keep=(),
                 graph=True):
        self.keep = tuple(keep)
        self.graph = graph
        self.precision = precision
        self.budget = budget
        self.quantum = quantum
        self.tiles = tiles or {2: 32, 3: 64}
        self.batches = {}
        self.active = {2: [], 3: []}
        self.bpp = dict(BYTES_PER_PARTICLE)  # admission charge per particle, per dimension
        self.measured = {2: False, 3: False}
        self.n_high = {2: 0, 3: 0}           # largest live particle count of each batch
        self.charge = {}                     # request id -> bytes charged
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_footprint = 0              # largest sph_batch_footprint total seen
        self.peak_active = 0
        self.pairs = 0
        self.particle_steps = 0
        self.pool = ThreadPoolExecutor(max_workers=2)

    def _batch(self, dim):
        b = self.batches.get(dim)
        if b is None:
            b = Batch(self.precision)
            b.set_tile(self.tiles[dim])
            b.set_graph(self.graph)
            self.batches[dim] = b
        return b

    def cost(self, r):
        return r.footprint(self.bpp[r.dim])

    def fits(self, r):
        return self.in_use + self.cost(r) <= self.budget

    def _measure(self, dim):
        """Bytes per live particle the batch of `dim` holds (sph_batch_footprint)."""
        b = self.batches[dim]
        _, n = b.size()
        if n > 0 and n >= self.n_high[dim]:
            self.n_high[dim] = n
            self.bpp[dim] = b.footprint() / n
            self.measured[dim] = True
        self.peak_footprint = max(self.peak_footprint,
                                  sum(x.footprint() for x in self.batches.values()))

    def admit(self, r, now):
        cost = self.cost(r)  # the charge the admission test used
        st = r.build()
        b = self._batch(r.dim)
        r.sid = b.add_state(st, periodic=r.periodic)
        r.admitted = now
        self.active[r.dim].append(r)
        self._measure(r.dim)  # the charge of later requests
        self.charge[r.rid] = cost
        self.in_use += cost
        self.peak_in_use = max(self.peak_in_use, self.in_use)

    def _advance(self, dim):
        """One quantum of the batch of `dim` (runs on a worker thread; the
        library call releases the interpreter lock)."""
        act = self.active[dim]
        b = self.batches[dim]
        b.init_rho0()  # the reference density of the requests that just joined
        k = min(self.quantum, min(r.steps - r.done_steps for r in act))
        b.step(k)
        dp, fp = b.step_pairs()
        return k, dp + fp

    def step_all(self, clock):
        """Advance every non-empty batch by one quantum, the 2D and 3D
        batches concurrently; returns the requests that finished."""
        done = []
        dims = [d for d in (2, 3) if self.active[d]]
        if len(dims) == 2:
            futs = {d: self.pool.submit(self._advance, d) for d in dims}
            res = {d: f.result() for d, f in futs.items()}
        else:
            res = {d: self._advance(d) for d in dims}
        for dim in dims:
            act = self.active[dim]
            b = self.batches[dim]
            k, pairs = res[dim]
            self.pairs += pairs
            self.particle_steps += k * sum(r.n for r in act)
            now = clock()
            keep = []
            for r in act:
                r.done_steps += k
                if r.done_steps >= r.steps:
                    r.result = summary(r.state, b, r.sid, self.keep)
                    r.finished = now
                    b.retire(r.sid)
                    self.in_use -= self.charge.pop(r.rid)
                    r.state = None  # host arrays are no longer needed
                    done.append(r)
                else:
                    keep.append(r)
            self.active[dim] = keep
        return done

    def run(self, requests, realtime=True, on_done=None):
        """Serve a request stream.  With realtime=True arrivals follow the
        wall clock (the server sleeps while idle); otherwise every request
        is available at once in arrival order.  Returns the requests with
        admission and completion times (seconds since the start)."""
        queue = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        qi = 0
        finished = []
        while qi < len(queue) or any(self.active.values()):
            now = clock()
            # FIFO admission: arrived and fits the budget (an oversize
            # request runs alone)
            while qi < len(queue) and (not realtime or queue[qi].arrival <= now):
                r = queue[qi]
                idle = not any(self.active.values())
                if not (self.fits(r) or idle):
                    break
                if not realtime:
                    r.arrival = now
                self.admit(r, now)
                qi += 1
            nact = sum(len(a) for a in self.active.values())
            self.peak_active = max(self.peak_active, nact)
            if nact == 0:
                if qi < len(queue):
                    time.sleep(max(0.0, min(0.05, queue[qi].arrival - clock())))
                continue
            for r in self.step_all(clock):
                finished.append(r)
                if on_done:
                    on_done(r)
        self.wall = clock()
        return finished

    def close(self):
        self.pool.shutdown()
        for b in self.batches.values():
            b.close()
        self.batches = {}


def latency_stats(finished):
    lat = np.array([r.latency for r in finished], dtype=np.float64)
    if len(lat) == 0:
        return {"p50_s": None, "p99_s": None, "mean_s": None}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean())}
