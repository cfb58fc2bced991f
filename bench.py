#!/usr/bin/env python3
"""Large-ring FP64 tracking benchmark on one MI355X.

Workload: the 6 GeV large ring (2,241 elements: radiating bends and
quadrupoles, sextupoles, octupoles, correctors, one aperture per cell, one RF
cavity), 262,144 particles from a seeded 6-D Gaussian (sigma = 10 x
(1e-4, 1e-5, 1e-5, 1e-6, 1e-3, 1e-3)), 1,000 turns. A step is one persistent
kernel launch tracking the whole beam for 100 turns with the particles
resident in HBM; 10 timed steps make the 1,000 turns.

value = particle-element passes per second, counting each particle until the
element where it is lost; roofline.achieved = the algorithmic FP64 FLOPs of a
launch (at_amd/flops.py) / the average launch time measured with HIP events
on the launch stream. cpu_baseline runs the reference integrator cores
(oracle/_ref, element-outer with OpenMP, serial loss scan as pyat/at.c) on a
bounded sample of the same workload on the host cores.

    python bench.py [--steps 10] [--warmup 1] [--turns-per-step 100]
    python bench.py --mode batch|serve|adjoint|monitors   (the other workloads, one line each)

The default line also carries ``serving``: one seeded trace of 96 requests
(error realisations resident on the device, N log-uniform on [2e3, 2e5],
turn counts scaled to {4, 16, 64} so that the offered loads 5, 15 and 50
requests/s bracket the capacity of both modes), served by continuous
batching and by one launch per request: p50/p99, throughput, the ratio at
equal p99, and whether every result is identical.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (= matrix) peak, TFLOP/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "track_kernel_large_ring_traffic.json")


def measured_traffic(particles: int, turns: int):
    """HBM bytes per launch from the FETCH_SIZE/WRITE_SIZE counter runs of the
    same launch shape (copy-calibrated), with the file they come from, or
    (None, None) when they were taken on another shape."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if t.get("particles") != particles or t.get("turns_per_launch") != turns:
        return None, None
    return t.get("bytes_per_launch"), os.path.relpath(TRAFFIC_FILE, ROOT)


METRIC = "particle-element passes/s FP64 + %FP64 peak; served p50/p99 latency vs plain"


def host_cores() -> dict:
    """The host as the CPU baseline sees it: nproc, the cores this process
    may use, the lease's thread budget (OMP_NUM_THREADS) and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
    return {"nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": omp or None,
            "threads": threads, "cpu_model": model}


def cpu_baseline(ring, beam_np, nturns: int):
    import numpy as np

    from oracle import pyoracle

    host = host_cores()
    threads = host["threads"]
    impl = "ref" if pyoracle.available("ref") else "oracle"
    r = np.asfortranarray(beam_np.copy())
    t0 = time.perf_counter()
    _, lm = pyoracle.atpass(ring, r, nturns, losses=True, impl=impl, nthreads=threads)
    dt = time.perf_counter() - t0
    from at_amd.flops import tracked_flops_and_passes

    _, passes = tracked_flops_and_passes(ring, nturns, 0, lm["islost"], lm["turn"], lm["elem"])
    return {
        "value": passes / dt,
        "unit": "particle-element passes/s",
        "cores": threads,
        "kind": "reference" if impl == "ref" else "port",
        "sample": f"{r.shape[1]} particles x {nturns} turns of the same ring and beam "
                  f"({passes:.3e} passes in {dt:.1f} s, OpenMP {threads} threads: every core "
                  f"the lease gives this process)",
        "host": host,
    }


def _pairs(n):
    from at_amd import device

    return [(device.Event(), device.Event()) for _ in range(n)]


def _same(a, b) -> bool:
    import numpy as np

    return bool(np.array_equal(a.to_host(), b.to_host(), equal_nan=True))


def bench_batch(args) -> dict:
    """Batched error realisations: ``--instances`` error-seeded copies of the
    large ring, each with a DA polar grid of U[1,000, 40,000] particles, all
    tracked for ``--batch-turns`` turns (1,024 by default) in ONE launch
    (launch shape from the autotuner), against the same launch without
    repacking and, with ``--per-instance``, the instances tracked one launch
    each. The realisations are the frozen orbit-corrected ones with 30 um rms
    shifts; ``--shift`` takes uncorrected realisations with that rms shift
    instead."""
    import numpy as np

    from at_amd import device, engine
    from at_amd.autotune import tune_batch
    from at_amd.batch import track_batch
    from at_amd.flops import tracked_flops_and_passes
    from at_amd.workloads import batched_da_workload

    work = batched_da_workload(ninst=args.instances, seed=7, shift=args.shift)
    lats = [l for l, _ in work]
    dls = [engine.DeviceLattice(l) for l in lats]
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    base = [device.DeviceArray.from_host(p.T) for _, p in work]
    for dl in dls:
        dl.check(energy, rest)
    T = args.batch_turns
    tune = tune_batch(dls, base, energy=energy, rest_energy=rest, use_cache=False,
                      repack=not args.no_repack, nturns=T)
    shape = {k: tune[k] for k in ("block_size", "ppt", "unroll")}
    print(f"batch: autotuned {shape}; {len(dls)} instances, {T} turns",
          file=sys.stderr, flush=True)
    stream = device.Stream()
    parts = [b.clone(stream) for b in base]
    (e0, e1), = _pairs(1)
    stream.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    lms = track_batch(dls, parts, T, energy=energy, rest_energy=rest, losses=True,
                      repack=not args.no_repack, stream=stream, **shape)
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    print(f"batch: launch {ms:.1f} ms", file=sys.stderr, flush=True)
    # the same launch without repacking between turn blocks
    plain_ms = plain_same = None
    if not args.no_repack and not args.no_compare:
        parts3 = [b.clone(stream) for b in base]
        (p0, p1), = _pairs(1)
        stream.synchronize()
        p0.record(stream)
        track_batch(dls, parts3, T, energy=energy, rest_energy=rest, losses=True,
                    stream=stream, **shape)
        p1.record(stream)
        stream.synchronize()
        plain_ms = p0.elapsed_ms(p1)
        print(f"batch: plain launch {plain_ms:.1f} ms", file=sys.stderr, flush=True)
        plain_same = all(_same(a, b) for a, b in zip(parts, parts3))
    flops = passes = 0.0
    for lat, lm in zip(lats, lms):
        fl, pa = tracked_flops_and_passes(lat, T, 0, lm.flag.to_host().astype(bool),
                                          lm.turn.to_host(), lm.elem.to_host())
        flops += fl
        passes += pa
    seq_ms = same = None
    if not args.no_compare and args.per_instance:
        parts2 = [b.clone(stream) for b in base]
        (s0, s1), = _pairs(1)
        stream.synchronize()
        s0.record(stream)
        for dl, p in zip(dls, parts2):
            dl.track_device(p.ptr, p.shape[0], T, energy=energy, rest_energy=rest,
                            stream=stream.handle)
        s1.record(stream)
        stream.synchronize()
        seq_ms = s0.elapsed_ms(s1)
        same = all(_same(a, b) for a, b in zip(parts, parts2))
    # every shape the autotuner scored, run as the whole launch
    tune_full = None
    if args.tune_check:
        tune_full = {}
        for name in sorted(tune["ms"]):
            b, rest_ = name[1:].split("p")
            p, u = rest_.split("u")
            parts4 = [x.clone(stream) for x in base]
            (t0_, t1_), = _pairs(1)
            stream.synchronize()
            t0_.record(stream)
            track_batch(dls, parts4, T, energy=energy, rest_energy=rest, losses=True,
                        repack=not args.no_repack, stream=stream, block_size=int(b), ppt=int(p),
                        unroll=int(u))
            t1_.record(stream)
            stream.synchronize()
            tune_full[name] = t0_.elapsed_ms(t1_)
            print(f"batch: shape {name}: {tune_full[name]:.1f} ms (model {tune['ms'][name]:.0f})",
                  file=sys.stderr, flush=True)
    achieved = flops / (ms * 1e-3) / 1e12
    ntot = sum(int(b.shape[0]) for b in base)
    return {
        "metric": "batched error realisations: particle-element passes/s (one launch)",
        "value": passes / (ms * 1e-3),
        "unit": "particle-element passes/s",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(dls)} error realisations of large_ring ({len(lats[0])} "
                               f"elements; rolls 100 urad, gradients 1e-3, shifts "
                               + ("30 um rms, orbit-corrected" if args.shift is None else
                                  f"{args.shift * 1e6:g} um rms, uncorrected")
                               + f"), DA polar grids of U[1000, 40000] "
                               f"particles ({ntot} total), {T} turns, one launch"
                               + ("" if args.no_repack else
                                  " (turn blocks with survivors repacked, one graph)"),
                   "launch_shape": shape, "autotune_model_ms": tune["ms"],
                   "autotune_pilot_turns": tune["pilot_turns"], "wall_s": wall},
        "roofline": {"bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None},
        "plain_launch_ms": plain_ms,
        "identical_to_plain_launch": plain_same,
        "one_launch_per_instance_ms": seq_ms,
        "speedup_vs_one_launch_per_instance": (seq_ms / ms) if seq_ms else None,
        "identical_to_one_launch_per_instance": same,
        "autotune_check": None if tune_full is None else {
            "full_launch_ms": tune_full,
            "chosen": f"b{shape['block_size']}p{shape['ppt']}u{shape['unroll']}",
            "best": min(tune_full, key=tune_full.get),
            "chosen_over_best": tune_full[f"b{shape['block_size']}p{shape['ppt']}u{shape['unroll']}"]
            / min(tune_full.values())},
    }


def _serving_setup(n_instances: int):
    from at_amd import engine
    from at_amd.workloads import corrected_realisations, large_ring

    base = large_ring()
    lats = corrected_realisations(base, n_instances)
    dls = [engine.DeviceLattice(l) for l in lats]
    for dl in dls:
        dl.check(base.energy, base.particle.rest_energy)
    return base, dls


def _serve_both(base, dls, trace, args, log=None) -> dict:
    """One trace through continuous batching and through one launch per
    request: throughput, p50/p99 and whether every result is identical."""
    import numpy as np

    from at_amd.serving import Server, alive_passes, serve_one_by_one

    energy, rest = base.energy, base.particle.rest_energy
    nel = len(base)
    srv = Server(dls, energy=energy, rest_energy=rest, turn_block=args.turn_block,
                 repack=not args.no_repack)
    res_b, st_b = srv.run(trace, log=log)
    res_p, st_p = serve_one_by_one(dls, trace, energy=energy, rest_energy=rest, log=log)
    passes = alive_passes(trace, res_b, lambda i: nel)
    same = all(np.array_equal(a.particles, b.particles, equal_nan=True)
               and np.array_equal(a.lost["islost"], b.lost["islost"])
               and np.array_equal(a.lost["turn"], b.lost["turn"])
               and np.array_equal(a.lost["elem"], b.lost["elem"])
               for a, b in zip(res_b, res_p))

    def side(st):
        return {"passes_per_s": passes / st["wall_s"], "requests_per_s": len(trace) / st["wall_s"],
                "p50_s": st["p50_s"], "p99_s": st["p99_s"], "wall_s": st["wall_s"],
                "launches": st["launches"]}

    return {"batched": side(st_b), "one_launch_per_request": side(st_p),
            "identical": bool(same), "passes": passes}


def serving_sweep(args) -> dict:
    """The served half of the metric: one seeded request trace (same requests
    at every rate, arrival times scaled) offered at rates that bracket the
    capacity of both modes, served by continuous batching and by one plain
    launch per request. ``at_equal_p99``: taking as latency bound the worst
    p99 one-launch-per-request reached, the highest throughput each mode
    delivered within it, and their ratio."""
    from at_amd.workloads import serving_trace

    rates = [float(r) for r in args.serve_rates.split(",")]
    turns = tuple(int(t) for t in args.serve_turns.split(","))
    base, dls = _serving_setup(args.serve_instances)
    # warm-up: a small request through each path
    warm = serving_trace(n_requests=2, rate=1e6, seed=99, n_instances=len(dls), nmin=64,
                         nmax=64, turns=(2,))
    _serve_both(base, dls, warm, args)
    points = []
    for rate in rates:
        trace = serving_trace(n_requests=args.serve_requests, rate=rate, seed=args.serve_seed,
                              n_instances=len(dls), turns=turns)
        print(f"serving: {len(trace)} requests at {rate:g}/s", file=sys.stderr, flush=True)
        pt = _serve_both(base, dls, trace, args)
        pt["offered_per_s"] = rate
        points.append(pt)
        b, p = pt["batched"], pt["one_launch_per_request"]
        print(f"serving: {rate:g}/s batched {b['requests_per_s']:.2f} req/s p99 {b['p99_s']:.3f} s;"
              f" plain {p['requests_per_s']:.2f} req/s p99 {p['p99_s']:.3f} s", file=sys.stderr,
              flush=True)
    bound = max(pt["one_launch_per_request"]["p99_s"] for pt in points)
    best = {m: max([pt[m]["passes_per_s"] for pt in points if pt[m]["p99_s"] <= bound],
                   default=None) for m in ("batched", "one_launch_per_request")}
    ratio = (best["batched"] / best["one_launch_per_request"]
             if best["batched"] and best["one_launch_per_request"] else None)
    return {
        "trace": {"requests": args.serve_requests, "seed": args.serve_seed,
                  "instances": len(dls), "particles": "log-uniform [2e3, 2e5]",
                  "turns": list(turns), "kinds": "half DA grids, half Gaussian beams",
                  "lattices": f"orbit-corrected error realisations (30 um shifts) of "
                              f"large_ring ({len(base)} elements), resident on the device",
                  "turn_block": args.turn_block},
        "points": points,
        "at_equal_p99": {"p99_bound_s": bound,
                         "batched_passes_per_s": best["batched"],
                         "one_launch_per_request_passes_per_s": best["one_launch_per_request"],
                         "ratio": ratio},
        "identical": all(pt["identical"] for pt in points),
    }


def bench_serve(args) -> dict:
    """Served mode at full request size: a Poisson trace of ``--requests``
    tracking and DA requests over ``--instances`` error realisations of the
    large ring, N log-uniform on [2e3, 2e5], turns in {256, 1024, 4096},
    served by continuous batching and by one launch per request."""
    from at_amd.workloads import serving_trace

    base, dls = _serving_setup(args.instances)
    trace = serving_trace(n_requests=args.requests, rate=args.rate, n_instances=len(dls))
    warm = serving_trace(n_requests=2, rate=1e6, seed=99, n_instances=len(dls), nmin=64,
                         nmax=64, turns=(2,))
    _serve_both(base, dls, warm, args)

    def log(r):
        print(f"request {r.rid} done at {r.done:.2f} s (latency {r.latency:.2f} s)",
              file=sys.stderr, flush=True)

    pt = _serve_both(base, dls, trace, args, log=log)
    b, p = pt["batched"], pt["one_launch_per_request"]
    return {
        "metric": "served tracking: particle-element passes/s, p50/p99 latency",
        "value": b["passes_per_s"],
        "unit": "particle-element passes/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": b["wall_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(trace)} requests, Poisson {args.rate}/s, "
                               f"{len(dls)} error realisations of large_ring ({len(base)} "
                               f"elements; rolls 100 urad, gradients 1e-3, shifts 30 um rms, "
                               f"orbit-corrected), N log-uniform [2e3, 2e5], turns in "
                               f"{{256, 1024, 4096}}",
                   "turn_block": args.turn_block,
                   "requests": [[q.npart, q.nturns, q.kind] for q in trace]},
        "batched": b, "one_launch_per_request": p,
        "throughput_ratio": p["wall_s"] / b["wall_s"], "identical_results": pt["identical"],
    }


def bench_shapes(args) -> dict:
    """The large-ring launch in every launch shape (workgroup size, particles
    per lane, slices per iteration): ``--particles`` particles, a 2-turn
    warm-up launch, then one timed launch of 20 turns; results do not depend
    on the shape, only the time does."""
    from at_amd import device, engine
    from at_amd.batch import LossMap
    from at_amd.flops import tracked_flops_and_passes
    from at_amd.workloads import gaussian_beam, large_ring

    ring = large_ring()
    n, T = args.particles, 20
    beam = gaussian_beam(n, seed=0, scale=10.0)
    dl = engine.DeviceLattice(ring)
    stream = device.Stream()
    kw = dict(energy=ring.energy, rest_energy=ring.particle.rest_energy, losses=True,
              stream=stream.handle)
    shapes = [(b, 1, u) for b in (64, 128, 256, 512) for u in (1, 2, 4)]
    shapes += [(b, 2, u) for b in (64, 128, 256, 512) for u in (1, 2)]
    out = []
    for sh in shapes:
        dl.set_launch_shape(*sh)
        r = device.DeviceArray.from_host(beam.T)
        lm0, lm = LossMap.empty(n), LossMap.empty(n)
        dl.track_device(r.ptr, n, 2, turn=0, loss_ptrs=lm0.ptrs(), **kw)
        (e0, e1), = _pairs(1)
        stream.synchronize()
        e0.record(stream)
        dl.track_device(r.ptr, n, T, turn=2, loss_ptrs=lm.ptrs(), **kw)
        e1.record(stream)
        stream.synchronize()
        ms = e0.elapsed_ms(e1)
        alive = ~lm0.flag.to_host().astype(bool)
        fl, _ = tracked_flops_and_passes(ring, T, 2, lm.flag.to_host().astype(bool)[alive],
                                         lm.turn.to_host()[alive], lm.elem.to_host()[alive])
        out.append({"block": sh[0], "ppt": sh[1], "unroll": sh[2], "ms_per_turn": ms / T,
                    "fp64_frac": fl / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS})
        print(f"shapes: {sh}: {ms / T:.3f} ms per turn", file=sys.stderr, flush=True)
        r.close()
    dl.close()
    best = min(out, key=lambda o: o["ms_per_turn"])
    return {"metric": "large-ring launch per launch shape: ms per turn", "unit": "ms",
            "value": best["ms_per_turn"], "higher_is_better": False, "n_gpus": 1,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"large_ring ({len(ring)} elements), {n} particles, turns "
                                   f"2-22, fast FP64 build"},
            "shapes": out, "best": best}


def bench_monitors(args) -> dict:
    """Monitor-heavy tracking: the large ring with a refpt before every
    element and at the end (2,242 copies of every particle per turn, the
    at.c output layout (6, N, R, T) kept in HBM), ``--particles`` particles,
    ``--steps`` launches of one turn each. The copies are 48 B per particle
    and refpt; the line reports their HBM bandwidth against 8 TB/s."""
    import numpy as np

    from at_amd import device, engine
    from at_amd.batch import LossMap
    from at_amd.flops import tracked_flops_and_passes
    from at_amd.workloads import gaussian_beam, large_ring

    ring = large_ring()
    nelem = len(ring)
    n = args.particles
    refs = np.arange(nelem + 1, dtype=np.uint32)
    R = refs.size
    beam = gaussian_beam(n, seed=0, scale=10.0)
    dl = engine.DeviceLattice(ring)
    stream = device.Stream()
    r_dev = device.DeviceArray.from_host(beam.T)
    rout = device.DeviceArray((R, n, 6), np.float64)
    K, W = args.steps, args.warmup
    lm = LossMap.empty(n)

    def launch(step):
        dl.track_device(r_dev.ptr, n, 1, refpts=refs, rout_ptr=rout.ptr, turn=step,
                        energy=ring.energy, rest_energy=ring.particle.rest_energy, losses=True,
                        loss_ptrs=lm.ptrs(), stream=stream.handle, strict=args.strict)

    for s in range(W):
        launch(s)
    stream.synchronize()
    ev = _pairs(K)
    flops = []
    alive_before = ~lm.flag.to_host().astype(bool)
    for k in range(K):
        ev[k][0].record(stream)
        launch(W + k)
        ev[k][1].record(stream)
        stream.synchronize()
        lost = lm.flag.to_host().astype(bool)
        fl, _ = tracked_flops_and_passes(ring, 1, W + k, (lost & alive_before)[alive_before],
                                         lm.turn.to_host()[alive_before],
                                         lm.elem.to_host()[alive_before])
        flops.append(fl)
        alive_before = ~lost
    ms = float(np.mean([a.elapsed_ms(b) for a, b in ev]))
    # per launch: state in and out, the refpt copies, the loss map zeroed
    # (57 B per particle) and written for lost particles
    nbytes = n * (48 + 48 + 57) + 48 * n * R
    gbs = nbytes / (ms * 1e-3) / 1e9
    tflops = float(np.mean(flops)) / (ms * 1e-3) / 1e12
    return {
        "metric": "refpt-heavy tracking: HBM GB/s of the turn-by-turn copies",
        "value": gbs, "unit": "GB/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"large_ring ({nelem} elements), {n} particles, refpts at "
                               f"every element and the end ({R} copies per particle per "
                               f"turn), 1 turn per launch, output (6, N, R, 1) in HBM",
                   "bytes_per_launch": nbytes},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                     "frac": gbs / 8000.0, "traffic": None},
        "fp64_tflops": tflops, "fp64_frac": tflops / PEAK_FP64_TFLOPS,
    }


def bench_adjoint(args) -> dict:
    """Differentiation: gradient of L = sum(x^2+px^2+y^2+py^2) after 64
    turns of the large ring (radiation off, cavity on) over 4,096 Gaussian
    particles, with respect to PolynomB[1] of every quadrupole and
    PolynomB[2] of every sextupole, checkpoints every 8 turns."""
    import numpy as np
    import torch

    from at_amd.adjoint import Problem, TrackingLoss, sweep_plan
    from at_amd.workloads import gaussian_beam, large_ring

    def _events(k):
        return [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                for _ in range(k)]

    ring = large_ring(radiation=False, cavity=True)
    params = []
    for i, e in enumerate(ring):
        if e.PassMethod.startswith("StrMPole") and e.MaxOrder in (1, 2):
            params.append((i, e.MaxOrder))
    prob = Problem(ring, params, checkpoint=8)
    n, T = args.adj_particles, 64
    beam = gaussian_beam(n, seed=5)
    dev = torch.device("cuda:0")
    parts = torch.from_numpy(np.ascontiguousarray(beam.T)).to(dev)
    theta0 = prob.initial_strengths()
    stream = torch.cuda.current_stream()

    def one():
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        loss = TrackingLoss.apply(prob, theta, parts, T)
        loss.backward()
        return loss, theta.grad

    one()
    torch.cuda.synchronize()
    K = args.adj_steps
    ev = _events(K)
    fw = _events(K)
    t0 = time.perf_counter()
    for k in range(K):
        ev[k][0].record(stream)
        theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        fw[k][0].record(stream)
        loss = TrackingLoss.apply(prob, theta, parts, T)
        fw[k][1].record(stream)
        loss.backward()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    fms = float(np.mean([a.elapsed_time(b) for a, b in fw]))
    g = sweep_plan(n, [8] * ((T + 7) // 8), 4 * torch.cuda.get_device_properties(
        dev).multi_processor_count)
    sweep = f"parallel, {g}-turn sub-blocks" if g else "sequential"
    seq_ms = None
    if not args.no_compare and g:
        # the same gradient with the blocks swept one after the other
        prob.parallel = False
        one()
        torch.cuda.synchronize()
        sq = _events(K)
        for k in range(K):
            sq[k][0].record(stream)
            one()
            sq[k][1].record(stream)
        torch.cuda.synchronize()
        seq_ms = float(np.mean([a.elapsed_time(b) for a, b in sq]))
        prob.parallel = None
    out = {
        "metric": "strength gradients: gradient evaluations/s (forward + adjoint)",
        "value": 1e3 / ms,
        "unit": "gradients/s",
        "n_gpus": 1, "steps": K, "warmup": 1, "ms_per_step": wall / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"large_ring radiation off, cavity on ({len(ring)} elements), "
                               f"{n} particles, {T} turns, {len(params)} strengths "
                               f"(PolynomB[1] of quadrupoles, PolynomB[2] of sextupoles), "
                               f"checkpoints every 8 turns",
                   "forward_ms": fms, "gradient_ms": ms, "sweep": sweep,
                   "sequential_sweep_gradient_ms": seq_ms,
                   "adjoint_over_forward": (ms - fms) / fms},
    }
    if not args.no_cpu:
        from oracle import pyoracle

        impl = "ref" if pyoracle.available("ref") else "oracle"
        threads = host_cores()["threads"]
        m = min(n, args.cpu_particles)
        r = np.asfortranarray(beam[:, :m].copy())
        t0 = time.perf_counter()
        pyoracle.atpass(ring, r, T, impl=impl, nthreads=threads)
        dt = (time.perf_counter() - t0) * (n / m)
        out["cpu_baseline"] = {
            "value": 1.0 / (2 * len(params) * dt), "unit": "gradients/s", "cores": threads,
            "kind": "reference" if impl == "ref" else "port",
            "sample": f"central differences need 2 x {len(params)} trackings of {n} particles x "
                      f"{T} turns; one tracking of {m} particles timed ({dt:.2f} s scaled to {n})",
        }
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=("track", "batch", "serve", "adjoint", "monitors",
                                       "shapes"),
                    default="track",
                    help="track (default): the large-ring tracking line with the serving sweep "
                         "and the CPU baseline; batch: many error realisations in one launch; "
                         "serve: continuous batching of a full-size request trace; adjoint: "
                         "strength gradients; monitors: refpts at every element (HBM "
                         "bandwidth of the copies); shapes: the large-ring launch in every "
                         "launch shape")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--batch-turns", type=int, default=1024)
    ap.add_argument("--shift", type=float, default=None,
                    help="batch: uncorrected error realisations with this rms quadrupole/"
                         "sextupole shift [m] (default: the orbit-corrected 30 um ones)")
    ap.add_argument("--no-repack", action="store_true",
                    help="batch/serve: one plain launch instead of repacked turn blocks")
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--rate", type=float, default=100.0)
    ap.add_argument("--turn-block", type=int, default=128)
    ap.add_argument("--no-serve", action="store_true",
                    help="track: leave out the serving sweep")
    ap.add_argument("--serve-requests", type=int, default=96)
    ap.add_argument("--serve-rates", default="5,15,50",
                    help="track: offered loads of the serving sweep [requests/s]")
    ap.add_argument("--serve-turns", default="4,16,64",
                    help="track: turn counts of the serving sweep's requests")
    ap.add_argument("--serve-instances", type=int, default=64)
    ap.add_argument("--serve-seed", type=int, default=11)
    ap.add_argument("--adj-particles", type=int, default=4096)
    ap.add_argument("--adj-steps", type=int, default=3)
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--repack", choices=("on", "off"), default="off",
                    help="track: repacked turn blocks (ATX_REPACK; gpupass's default is off)")
    ap.add_argument("--tune-check", action="store_true",
                    help="batch: also run every shape the autotuner scored as the whole launch")
    ap.add_argument("--per-instance", action="store_true",
                    help="batch: also time one launch per instance (about 6 minutes at 1,024 "
                         "turns)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--turns-per-step", type=int, default=100)
    ap.add_argument("--particles", type=int, default=262144)
    ap.add_argument("--strict", action="store_true", help="reference operation order build")
    ap.add_argument("--cpu-particles", type=int, default=32768)
    ap.add_argument("--cpu-turns", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 times): the "
                         "calibration kernel for the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (particles "
              f"and lattice instances shard across GPUs with no exchange during tracking; "
              f"the bench measures one GPU's share)", file=sys.stderr)
        return 2

    import numpy as np

    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe * 131072, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "dispatches": 3}))
        return 0

    from at_amd import device, engine
    from at_amd.batch import LossMap
    from at_amd.flops import lattice_flops, tracked_flops_and_passes
    from at_amd.workloads import gaussian_beam, large_ring

    engine.set_device(0)
    if args.mode != "track":
        fn = {"batch": bench_batch, "serve": bench_serve, "adjoint": bench_adjoint,
              "monitors": bench_monitors, "shapes": bench_shapes}[args.mode]
        print(json.dumps(fn(args)), flush=True)
        return 0
    ring = large_ring()
    nelem = len(ring)
    n = args.particles
    T = args.turns_per_step
    beam = gaussian_beam(n, seed=0, scale=10.0)
    dl = engine.DeviceLattice(ring)
    stream = device.Stream()

    # particles resident in HBM: (N, 6) row-major == (6, N) F-order
    r_dev = device.DeviceArray.from_host(beam.T)
    K, W = args.steps, args.warmup
    loss_sets = [LossMap.empty(n) for _ in range(W + K)]

    repack = args.repack == "on"

    def launch(step: int):
        dl.track_device(r_dev.ptr, n, T, turn=step * T, energy=ring.energy,
                        rest_energy=ring.particle.rest_energy, losses=True,
                        loss_ptrs=loss_sets[step].ptrs(), stream=stream.handle,
                        strict=args.strict, repack=repack)

    print(f"track: {W} warm-up and {K} timed launches of {T} turns, {n} particles",
          file=sys.stderr, flush=True)
    for s in range(W):
        launch(s)
    stream.synchronize()

    events = _pairs(K)
    t0 = time.perf_counter()
    for k in range(K):
        events[k][0].record(stream)
        launch(W + k)
        events[k][1].record(stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_ms(b) for a, b in events]

    # algorithmic work of each timed launch, from its loss map
    lost_before = np.zeros(n, bool)
    for s in range(W):
        lost_before |= loss_sets[s].flag.to_host().astype(bool)
    flops_tot = passes_tot = 0.0
    flops_per_launch = []
    for k in range(K):
        lm = loss_sets[W + k]
        lf, lt, le = lm.flag.to_host(), lm.turn.to_host(), lm.elem.to_host()
        lost = lf.astype(bool) & ~lost_before
        alive = ~lost_before
        fl, pa = tracked_flops_and_passes(ring, T, (W + k) * T, lost[alive], lt[alive], le[alive])
        flops_per_launch.append(fl)
        flops_tot += fl
        passes_tot += pa
        lost_before |= lf.astype(bool)
    survivors = int((~lost_before).sum())
    mean_ms = float(np.mean(kernel_ms))
    achieved = float(np.mean(flops_per_launch)) / (mean_ms * 1e-3) / 1e12
    traffic, traffic_src = (None, None) if args.strict else measured_traffic(n, T)
    print(f"track: {mean_ms:.1f} ms per launch, {achieved:.2f} TFLOP/s", file=sys.stderr,
          flush=True)
    line = {
        "metric": METRIC,
        "value": passes_tot / wall,
        "unit": "particle-element passes/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"large_ring: 6 GeV, {nelem} elements (radiation on, 1 RF cavity, "
                        f"apertures), {n} particles, {K * T} turns ({K} steps x {T} turns), "
                        f"{'strict' if args.strict else 'fast'} FP64 build",
            "particles": n,
            "elements": nelem,
            "turns": K * T,
            "survivors": survivors,
            "flop_per_particle_turn": float(lattice_flops(ring).sum()),
            "repacked_turn_blocks": repack,
        },
        "roofline": {
            "bound": "fp64_valu",
            "pipe": "FP64 vector ALU (no MFMA: the path is not a contraction); MI355X FP64 "
                    "vector peak = FP64 matrix peak",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "traffic_source": (f"{traffic_src}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of "
                               f"this launch shape, copy-calibrated; not measured in this run"
                               if traffic_src else None),
            "kernel_ms_mean": mean_ms,
            "flops_per_launch": float(np.mean(flops_per_launch)),
        },
    }
    if not args.no_serve:
        line["serving"] = serving_sweep(args)
    if not args.no_cpu:
        sample = beam[:, : args.cpu_particles]
        print(f"cpu baseline: {sample.shape[1]} particles x {args.cpu_turns} turns",
              file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(ring, sample, args.cpu_turns)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
