"""Benchmark: pair-interactions/s of the 3D WCSPH dam break on one MI355X.

Workload (default): 3D dam break, dr = 0.0079, 4,096,766 fluid + 1,655,459
wall particles, neighbour radius R = 4.8 dr (the Wendland support 2 hs,
hs = 2.4 dr), FP32-mixed pair arithmetic with FP64 state.  A step is one
fused device step: cell-list build, density sum, EOS, internal force with
artificial viscosity, symplectic Euler kick and drift.

  value = 2 * (accepted ordered pairs per step) * steps / wall time
          (two pair passes per step: density, force), inputs resident in HBM.

The same workload in plain FP64 (the drop-in path) is timed beside it, and
the FP64 CPU oracle (a C port of the reference algorithm, OpenMP over
particles) is timed on two steps of the same workload on the host cores.

Beside the headline the same JSON line carries:
  accuracy_vs_fp64  mixed against FP64 on the device after 100 and 1000 steps
                    (relative density, velocity and total-energy error);
  tg_batch          256 periodic 2D Taylor-Green boxes (N log-uniform in
                    [2k, 60k], Re uniform in [50, 1000]) advanced 500 steps as
                    one batch, mixed and FP64, with the kinetic-energy decay
                    against the FP64 run and the closed form;
  serving           a seeded Poisson stream of 2D/3D dam-break and
                    Taylor-Green requests (1k-500k particles, 200-2000 steps)
                    served by continuous batching in the mixed mode: p50/p99
                    completion latency and pair-interactions/s, beside the same
                    requests run one at a time in plain FP64.
  probe_sums        SmoothedParticles.sum at 65536 seeded points of the fluid's
                    box (Shepard and velocity sums), per call with host buffers.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64]
       [--no-fp64] [--no-cpu] [--no-accuracy] [--no-batch] [--no-serve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "smoothedparticles.jl_amd"))

# Published peaks of one MI355X.  The FP32 vector figure counts packed FP32
# (two operations per lane per instruction); this build issues no packed
# FP32, so the rate its kernels can reach is the non-packed one, 78.6
# TFLOP/s, the same as FP64.
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP32_NONPACKED_TFLOPS = 78.6
PEAK_FP64_TFLOPS = 78.6

# Algorithmic work of the two pair passes of a step.
#   density pass: the distance tests it performs (one bounding-box test per
#     candidate per tile, one pair test per kept candidate per target; about
#     9 flops each -- 3 sub or fma, squares and sums, compares), every
#     accepted pair adds m W (about 15 flops); bytes: positions read and rho
#     written once in sorted order (the accepted-candidate masks it writes
#     are scratch, not algorithmic).  The reference's search tests every
#     candidate of the 27 (9) cells; cand_all reports that count beside.
#   force pass: only the accepted pairs of fluid targets (the masks of the
#     density pass say which), about 45 flops each (pressure gradient, Monaghan
#     viscosity, FP64 accumulation); bytes: x, v, rho, P/rho^2, type read and a
#     written once.
FLOPS_PER_CANDIDATE = 9.0
FLOPS_PER_DENSITY_PAIR = 15.0
FLOPS_PER_FORCE_PAIR = 45.0
DENSITY_BYTES_PER_PARTICLE = {"mixed": 32.0, "fp64": 32.0}
FORCE_BYTES_PER_PARTICLE = {"mixed": 92.0, "fp64": 92.0}

# Density summation with the per-particle reference density rho0_p set by
# one find_rho0! pass at t = 0 (tests/test_collision_2d.jl:123-126): the EOS
# is P = c^2 (rho - rho0_p), so kernel truncation at walls and the free
# surface does not turn into pressure.
SCHEME = "wcsph_se_rho0"

WORKLOADS = {
    "dambreak3d_4M": dict(dim=3, dr=0.0079, R_over_dr=4.8,
                          desc="3D WCSPH dam break, dr=0.0079, R=4.8dr"),
    "dambreak3d_1M": dict(dim=3, dr=0.0125, R_over_dr=4.8,
                          desc="3D WCSPH dam break, dr=0.0125, R=4.8dr"),
    "dambreak2d": dict(dim=2, dr=0.02, R_over_dr=4.8,
                       desc="2D WCSPH dam break, dr=0.02, R=4.8dr"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(st, precision, steps, warmup, autotune=True):
    import smoothedparticles_hip as sp

    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, precision)
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    sp.init_reference_density(g, **p)
    T = g.autotune(**p) if autotune else None
    if warmup:
        sp.step(g, warmup, scheme=SCHEME, **p)
    g.profile(True)
    t0 = time.perf_counter()
    pairs = sp.step(g, steps, scheme=SCHEME, **p)  # returns after a stream sync
    t1 = time.perf_counter()
    dens_pairs, force_pairs = g.step_pairs()
    tests = g.step_tests()
    force_ms, force_n = g.profile_read(sp._native.SPH_KERNEL_FORCE)
    dens_ms, dens_n = g.profile_read(sp._native.SPH_KERNEL_DENSITY)
    build_ms, build_n = g.profile_read(sp._native.SPH_KERNEL_BUILD)
    g.profile(False)
    sp.create_cell_list(g)
    cand_all, cand_fluid = g.candidates()
    probe = probe_leg(sp, g, st, p)
    n = len(g)
    g.close()
    el = t1 - t0
    return dict(
        elapsed=el, pairs=pairs, value=2.0 * pairs / el, n=n, tile=T,
        force_ms=force_ms / max(force_n, 1), density_ms=dens_ms / max(dens_n, 1),
        build_ms=build_ms / max(build_n, 1), force_pairs_per_pass=force_pairs / steps,
        density_pairs_per_pass=dens_pairs / steps, tests_per_pass=tests / steps,
        cand_fluid=cand_fluid, cand_all=cand_all, probe=probe,
    )


PROBES = 1 << 16


def probe_points(st):
    import numpy as np

    rng = np.random.default_rng(1)
    xf = st.x[st.type == 0.0]
    lo, hi = xf.min(axis=0), xf.max(axis=0)
    return lo + rng.random((PROBES, 3)) * (hi - lo)


def probe_leg(sp, g, st, p, reps=5):
    """SmoothedParticles.sum at PROBES seeded points in the fluid's bounding
    box (Shepard denominator and velocity sums, fluid only), timed per call
    with host buffers: staging, both PCIe copies and the kernel."""
    P = probe_points(st)
    sp.probe_sums(g, P, fluid_only=True, **p)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = sp.probe_sums(g, P, fluid_only=True, **p)
    el = (time.perf_counter() - t0) / reps
    filled = out[:, 0] > 0
    return {"probes": PROBES, "call_ms": 1e3 * el, "probes_per_s": PROBES / el,
            "probes_with_fluid": int(filled.sum()),
            "timing": "per call with host buffers (staging, PCIe both ways, kernel)"}


def accuracy(st, checkpoints=(100, 1000)):
    """Mixed against plain FP64 on the device, same initial state: max
    relative density error, velocity error (relative to max |v|) and
    relative total-energy error at each checkpoint."""
    import numpy as np

    import smoothedparticles_hip as sp

    p = st.params()
    sysm = {}
    for prec in ("fp64", "mixed"):
        g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, prec)
        g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
        sp.init_reference_density(g, **p)
        sysm[prec] = g
    out, done = {}, 0
    for c in checkpoints:
        for g in sysm.values():
            sp.step(g, c - done, scheme=SCHEME, **p)
        done = c
        a, b = sysm["mixed"], sysm["fp64"]
        if len(a) != len(b):
            raise RuntimeError(f"particle counts differ: mixed {len(a)}, fp64 {len(b)}")
        rho_a, rho_b = a.get("rho"), b.get("rho")
        v_a, v_b = a.get("v"), b.get("v")
        e_a, e_b = sp.energy(a, "wcsph_rho0", **p), sp.energy(b, "wcsph_rho0", **p)
        out[str(c)] = dict(particles=len(a), rho_rel=float(np.max(np.abs(rho_a / rho_b - 1))),
                           v_rel=float(np.max(np.abs(v_a - v_b)) / np.max(np.abs(v_b))),
                           energy_rel=float(abs(e_a / e_b - 1)))
        log(f"step {c}: {out[str(c)]}")
    for g in sysm.values():
        g.close()
    return out


def tg_batch(steps=500, count=256, seed=0):
    """The batched periodic Taylor-Green workload, mixed and FP64."""
    import math

    import numpy as np

    import smoothedparticles_hip as sp
    from smoothedparticles_hip import states

    spec = states.taylor_green_batch(count, seed=seed)
    sts = [states.taylor_green(n, Re) for n, Re in spec]
    # warm the kernels up on a throwaway batch of a few boxes
    wb = sp.Batch("mixed")
    for st in sts[:4]:
        wb.add_state(st, periodic=True)
    wb.init_rho0()
    wb.step(2)
    wb.close()
    res = {}
    ke = {}
    for prec in ("mixed", "fp64"):
        b = sp.Batch(prec)
        ids = [b.add_state(st, periodic=True) for st in sts]
        b.init_rho0()
        T = b.autotune()
        t0 = time.perf_counter()
        b.step(steps)
        el = time.perf_counter() - t0
        dp, fp = b.step_pairs()
        ke[prec] = np.array([states.taylor_green_kinetic(st, b.get(i, "v"))
                             for st, i in zip(sts, ids)])
        res[prec] = dict(value=(dp + fp) / el, ms_per_step=1e3 * el / steps, tile_targets=T,
                         footprint_GB=b.footprint() / 1e9)
        b.close()
        log(f"tg_batch {prec}: {res[prec]}")
    ke0 = np.array([states.taylor_green_kinetic(st, st.v) for st in sts])
    exact = np.array([math.exp(-16 * math.pi ** 2 * st.nu * steps * st.dt) for st in sts])
    rel_fp64 = np.abs(ke["mixed"] / ke["fp64"] - 1)
    rel_exact = np.abs(ke["fp64"] / ke0 / exact - 1)
    return {
        "workload": f"{count} periodic 2D Taylor-Green unit boxes, N log-uniform [2k, 60k] "
                    f"(total {sum(len(st.x) for st in sts)}), Re uniform [50, 1000], R=4.8dr, "
                    f"{steps} steps, one launch per stage for all boxes",
        "value": res["mixed"]["value"], "unit": "pair-interactions/s",
        "ms_per_step": res["mixed"]["ms_per_step"], "tile_targets": res["mixed"]["tile_targets"],
        "fp64_value": res["fp64"]["value"], "fp64_ms_per_step": res["fp64"]["ms_per_step"],
        "mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
        "ke_rel_vs_fp64_max": float(rel_fp64.max()),
        "ke_rel_vs_closed_form_median": float(np.median(rel_exact)),
        "ke_rel_vs_closed_form_max": float(rel_exact.max()),
    }


def serve(count=48, rate=2.5, seed=0, budget_gb=32.0):
    """The served request stream, mixed with continuous batching, and the
    same requests one at a time in plain FP64 (single-system handles)."""
    import numpy as np

    import smoothedparticles_hip as sp
    from smoothedparticles_hip import serving

    reqs = serving.make_requests(count, rate, seed=seed)
    t0 = time.perf_counter()
    for r in reqs:
        r.build()
    log(f"serving: {count} request states built in {time.perf_counter() - t0:.1f} s")
    total_particles = sum(r.n for r in reqs)
    srv = serving.Server("mixed", budget=int(budget_gb * (1 << 30)), quantum=50)
    fin = srv.run(reqs, realtime=True)
    lat = serving.latency_stats(fin)
    out = {
        "workload": f"{count} requests, Poisson rate {rate}/s, kinds {list(serving.KINDS)}, "
                    f"1k-500k particles (total {total_particles}), 200-2000 steps, seed {seed}",
        "value": srv.pairs / srv.wall, "unit": "pair-interactions/s",
        "wall_s": srv.wall, "p50_s": lat["p50_s"], "p99_s": lat["p99_s"],
        "mean_latency_s": lat["mean_s"], "peak_concurrent": srv.peak_active,
        "budget_GB": srv.budget / 1e9, "peak_admitted_GB": srv.peak_in_use / 1e9,
        "peak_device_footprint_GB": srv.peak_footprint / 1e9, "step_graphs": srv.graph,
    }
    srv.close()
    log(f"serving mixed: {out}")
    # sequential plain FP64: service time of each request alone, then the
    # same arrivals through a FIFO single server
    service, pairs_tot = [], 0
    for r in sorted(reqs, key=lambda r: r.rid):
        st = r.build()
        p = st.params()
        t1 = time.perf_counter()
        if r.periodic:
            g = sp.Batch("fp64")
            sid = g.add_state(st, periodic=True)
            g.init_rho0()
            g.step(r.steps)
            dp, fp = g.step_pairs()
        else:
            g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "fp64")
            g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
            sp.init_reference_density(g, **p)
            sp.step(g, r.steps, scheme=SCHEME, **p)
            dp, fp = g.step_pairs()
        g.close()
        service.append(time.perf_counter() - t1)
        pairs_tot += dp + fp
        r.state = None
    arrivals = [r.arrival for r in sorted(reqs, key=lambda r: r.rid)]
    finish, lat_seq = 0.0, []
    for a, sv in zip(arrivals, service):
        finish = max(finish, a) + sv
        lat_seq.append(finish - a)
    out["fp64_sequential_value"] = pairs_tot / sum(service)
    out["fp64_sequential_p50_s"] = float(np.percentile(lat_seq, 50))
    out["fp64_sequential_p99_s"] = float(np.percentile(lat_seq, 99))
    out["speedup_vs_fp64_sequential"] = out["value"] / out["fp64_sequential_value"]
    log(f"serving fp64 sequential: {out['fp64_sequential_value']:.4g} pair-int/s, "
        f"p50 {out['fp64_sequential_p50_s']:.2f} s, p99 {out['fp64_sequential_p99_s']:.2f} s")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_c0_rates():
    """The oracle on the 2D dam break (dr = 0.02, 6514 particles, the
    reference's CPU-runnable case): pair-interactions/s of 200 full steps on
    one thread and on every host thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from smoothedparticles_hip import states

    st = states.dam_break(2, 0.02)
    out = {}
    for label, threads in (("1_thread", 1), ("all_threads", O.num_threads())):
        O.set_threads(threads)
        o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
        o.load(st.x, st.v, st.rho, st.P, st.type)
        o.create_cell_list()
        o.density_sum_rho0(True, **st.params())
        t0 = time.perf_counter()
        pairs = o.step_wcsph(200, use_rho0=True, **st.params())
        out[label] = {"value": 2.0 * pairs / (time.perf_counter() - t0), "threads": threads}
    O.set_threads(out["all_threads"]["threads"])
    return out


def cpu_baseline(st):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
    o.load(st.x, st.v, st.rho, st.P, st.type)
    o.create_cell_list()
    o.density_sum_rho0(True, **st.params())
    nsteps = 2  # about 10 s on 16 host threads
    t0 = time.perf_counter()
    pairs = o.step_wcsph(nsteps, use_rho0=True, **st.params())
    el = time.perf_counter() - t0
    c0 = cpu_c0_rates()
    # probe sums: the oracle's serial loop over the first 4096 of the bench's probes
    o.create_cell_list()
    P = probe_points(st)[:4096]
    t1 = time.perf_counter()
    o.probe_sum(P, st.params()["m"], True)
    el_p = time.perf_counter() - t1
    return dict(value=2.0 * pairs / el, unit="pair-interactions/s", cores=O.num_threads(),
                cpu_model=cpu_model(), host_threads=os.cpu_count(),
                c0_2d_dambreak_200_steps=c0,
                probe_sums_per_s=len(P) / el_p, probe_cores=1,
                probe_sample=f"{len(P)} of the probe leg's points after the steps, serial C loop, "
                             f"{el_p:.2f} s",
                kind="port",
                sample=f"{nsteps} full steps (cell list, density, EOS, force, integrate) of the "
                       f"same {len(st.x)}-particle workload by the FP64 C port of the reference "
                       f"algorithm, OpenMP over particles; {el:.1f} s")


def accuracy_vs_oracle(dr=0.02, steps=100):
    """The 3D dam break in the mixed form against the FP64 oracle at the
    largest size the oracle runs in minutes (dr = 0.02, about 0.5M particles):
    relative density, velocity and total-energy error every 25 steps."""
    import numpy as np

    import smoothedparticles_hip as sp
    from smoothedparticles_hip import states

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    st = states.dam_break(3, dr)
    p = st.params()
    g = sp.ParticleSystem(sp.Box(*st.bbox_lo, *st.bbox_hi), st.R, "mixed")
    g.add_particles(st.x, v=st.v, rho=st.rho, P=st.P, type=st.type)
    sp.init_reference_density(g, **p)
    o = O.OracleSystem(st.R, st.bbox_lo, st.bbox_hi)
    o.load(st.x, st.v, st.rho, st.P, st.type)
    o.create_cell_list()
    o.density_sum_rho0(True, **p)
    out, done = {}, 0
    for c in range(25, steps + 1, 25):
        sp.step(g, c - done, scheme=SCHEME, **p)
        t0 = time.perf_counter()
        o.step_wcsph(c - done, use_rho0=True, **p)
        done = c
        vo = o.field("v")
        out[str(c)] = dict(
            rho_rel=float(np.max(np.abs(g.get("rho") / o.field("rho") - 1))),
            v_rel=float(np.max(np.abs(g.get("v") - vo)) / np.max(np.abs(vo))),
            energy_rel=float(abs(sp.energy(g, "wcsph_rho0", **p) / o.energy("wcsph_rho0", **p) - 1)),
            particles=len(g), oracle_s=time.perf_counter() - t0)
        log(f"vs oracle, step {c}: {out[str(c)]}")
    return {"workload": f"3D dam break dr={dr}, {len(st.x)} particles, mixed device path vs FP64 "
                        f"oracle (C port of the reference algorithm), {steps} steps",
            "errors": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dambreak3d_4M", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the mixed-vs-FP64 errors at steps 100 and 1000")
    ap.add_argument("--no-batch", action="store_true", help="skip the Taylor-Green batch")
    ap.add_argument("--no-serve", action="store_true", help="skip the served request stream")
    ap.add_argument("--accuracy-oracle", action="store_true",
                    help="also run the mixed 3D dam break (dr=0.02) against the FP64 oracle "
                         "for 100 steps (several minutes of host time)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(each system runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from smoothedparticles_hip import states

    w = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    st = states.dam_break(w["dim"], w["dr"], R_over_dr=w["R_over_dr"])
    log(f"state: {len(st.x)} particles ({st.n_fluid} fluid) in {time.perf_counter() - t0:.1f} s")

    res = run_gpu(st, args.precision, args.steps, args.warmup)
    log(f"{args.precision}: {res['value']:.4g} pair-int/s, {1e3 * res['elapsed'] / args.steps:.2f} ms/step, "
        f"force {res['force_ms']:.3f} ms, density {res['density_ms']:.3f} ms, build {res['build_ms']:.3f} ms, T={res['tile']}")
    other = None
    if not args.no_fp64 and args.precision == "mixed":
        other = run_gpu(st, "fp64", args.steps, args.warmup)
        log(f"fp64: {other['value']:.4g} pair-int/s, force {other['force_ms']:.3f} ms, T={other['tile']}")

    prec = args.precision
    peak = PEAK_FP32_TFLOPS if prec == "mixed" else PEAK_FP64_TFLOPS

    def roofline(kernel):
        if kernel == "density":
            flops = (FLOPS_PER_CANDIDATE * res["tests_per_pass"]
                     + FLOPS_PER_DENSITY_PAIR * res["density_pairs_per_pass"])
            ms, bpp = res["density_ms"], DENSITY_BYTES_PER_PARTICLE[prec]
            name = "density pass (candidate test, pair masks, rho = sum m W)"
        else:
            flops = FLOPS_PER_FORCE_PAIR * res["force_pairs_per_pass"]
            ms, bpp = res["force_ms"], FORCE_BYTES_PER_PARTICLE[prec]
            name = "force pass (pressure gradient + artificial viscosity over masked pairs)"
        achieved = flops / (ms * 1e-3) / 1e12
        hbm_gbs = bpp * res["n"] / (ms * 1e-3) / 1e9
        traffic, tsrc = None, None
        tpath = os.path.join(ROOT, "profiles", f"{kernel}_pass_{args.workload}_{prec}_traffic.json")
        if os.path.exists(tpath):
            tj = json.load(open(tpath))
            traffic = tj.get("bytes_per_launch")
            tsrc = (f"profiles/{os.path.basename(tpath)}: rocprofv3 --pmc FETCH_SIZE and "
                    f"WRITE_SIZE of this kernel, measured {tj.get('measured', 'earlier')}, "
                    "not in this run")
        return {
            "bound": "valu",
            "ceiling": "vector ALU: the pair kernels issue no MFMA (the field's 'mfma' does not "
                       "apply); peak is the published "
                       f"{'packed FP32' if prec == 'mixed' else 'FP64'} vector rate",
            "kernel": name,
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "frac_vs_nonpacked_fp32": (achieved / PEAK_FP32_NONPACKED_TFLOPS
                                       if prec == "mixed" else None),
            "traffic": traffic,
            "traffic_source": tsrc,
            "flops_per_launch": flops,
            "hbm_algorithmic_GBps": hbm_gbs,
            "hbm_frac": hbm_gbs / PEAK_HBM_GBS,
            "avg_launch_ms": ms,
        }

    dominant = "density" if res["density_ms"] >= res["force_ms"] else "force"
    secondary = "force" if dominant == "density" else "density"

    out = {
        "metric": "pair-interactions/s, FP32-mixed vs FP64 plain at stated error; p50/p99 latency",
        "value": res["value"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed" if prec == "mixed" else "f64",
        "data": "synthetic (seeded lattice dam break)",
        "config": {
            "workload": w["desc"] + f", {res['n']} particles ({st.n_fluid} fluid), "
                        f"{'FP32-mixed pair arithmetic, FP64 state' if prec == 'mixed' else 'plain FP64'}",
            "particles": res["n"],
            "neighbour_radius": st.R,
            "pairs_per_step": res["pairs"] / args.steps,
            "tile_targets": res["tile"],
            "distance_tests_per_density_pass": res["tests_per_pass"],
            "reference_candidates_per_pass": res["cand_all"],
            "parallelism": "single GPU",
        },
        "roofline": roofline(dominant),
        "roofline_" + secondary: roofline(secondary),
    }
    out["probe_sums"] = res["probe"]
    if other is not None:
        out["fp64_value"] = other["value"]
        out["mixed_over_fp64"] = res["value"] / other["value"]
        out["fp64_ms_per_step"] = 1e3 * other["elapsed"] / args.steps
    if not args.no_accuracy:
        out["accuracy_vs_fp64"] = accuracy(st)
    if args.accuracy_oracle:
        out["accuracy_vs_oracle"] = accuracy_vs_oracle()
    if not args.no_batch:
        out["tg_batch"] = tg_batch()
    if not args.no_serve:
        out["serving"] = serve()
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(st)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
