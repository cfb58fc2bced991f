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
METRIC = "particle-element passes/s FP64 + %FP64 peak; served p50/p99 latency vs plain"


def cpu_baseline(ring, beam_np, nturns: int, threads: int):
    import numpy as np

    from oracle import pyoracle

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
                  f"({passes:.3e} passes in {dt:.1f} s, OpenMP {threads} threads)",
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--turns-per-step", type=int, default=100)
    ap.add_argument("--particles", type=int, default=262144)
    ap.add_argument("--strict", action="store_true", help="reference operation order build")
    ap.add_argument("--cpu-particles", type=int, default=32768)
    ap.add_argument("--cpu-turns", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (3 times): the "
                         "calibration kernel for the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              f"(particles shard with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_probe:
        src = torch.ones(args.copy_probe * 131072, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "dispatches": 3}))
        return 0

    from at_amd import engine
    from at_amd.flops import lattice_flops, tracked_flops_and_passes
    from at_amd.workloads import gaussian_beam, large_ring

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    engine.set_device(0)
    ring = large_ring()
    nelem = len(ring)
    n = args.particles
    T = args.turns_per_step
    beam = gaussian_beam(n, seed=0, scale=10.0)
    dl = engine.DeviceLattice(ring)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    # particles resident in HBM: (N, 6) row-major == (6, N) F-order
    r_dev = torch.from_numpy(np.ascontiguousarray(beam.T)).to(dev)
    K, W = args.steps, args.warmup
    loss_sets = []
    for _ in range(W + K):
        loss_sets.append((torch.zeros(n, dtype=torch.uint8, device=dev),
                          torch.zeros(n, dtype=torch.int32, device=dev),
                          torch.zeros(n, dtype=torch.int32, device=dev),
                          torch.zeros((n, 6), dtype=torch.float64, device=dev)))

    def launch(step: int):
        lf, lt, le, lc = loss_sets[step]
        dl.track_device(r_dev.data_ptr(), n, T, turn=step * T, energy=ring.energy,
                        rest_energy=ring.particle.rest_energy, losses=True,
                        loss_ptrs=(lf.data_ptr(), lt.data_ptr(), le.data_ptr(), lc.data_ptr()),
                        stream=sptr, strict=args.strict)

    for s in range(W):
        launch(s)
    torch.cuda.synchronize()

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        events[k][0].record(stream)
        launch(W + k)
        events[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in events]

    # algorithmic work of each timed launch, from its loss map
    lost_before = np.zeros(n, bool)
    for s in range(W):
        lost_before |= loss_sets[s][0].cpu().numpy().astype(bool)
    flops_tot = passes_tot = 0.0
    flops_per_launch = []
    for k in range(K):
        lf, lt, le, _ = (x.cpu().numpy() for x in loss_sets[W + k])
        lost = lf.astype(bool) & ~lost_before
        alive = ~lost_before
        fl, pa = tracked_flops_and_passes(ring, T, (W + k) * T, lost[alive],
                                          lt.view(np.uint32)[alive], le.view(np.uint32)[alive])
        flops_per_launch.append(fl)
        flops_tot += fl
        passes_tot += pa
        lost_before |= lf.astype(bool)
    survivors = int((~lost_before).sum())
    mean_ms = float(np.mean(kernel_ms))
    achieved = float(np.mean(flops_per_launch)) / (mean_ms * 1e-3) / 1e12
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
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "fp64 vector (VALU); MI355X FP64 vector peak = FP64 matrix peak",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": None,
            "kernel_ms_mean": mean_ms,
            "flops_per_launch": float(np.mean(flops_per_launch)),
        },
    }
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        sample = beam[:, : args.cpu_particles]
        line["cpu_baseline"] = cpu_baseline(ring, sample, args.cpu_turns, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
