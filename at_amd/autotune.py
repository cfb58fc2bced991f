"""Launch-shape autotuner for batched tracking.

Picks the workgroup size, the particles per lane and the slices per loop
iteration of one batched launch (include/atx.h atx_launch_shape; results do
not depend on the shape). A launch of many turns spends its first turns on
the full beam and the rest on the survivors, which the repacked launch packs
into fewer workgroups; the two regimes favour different shapes. So each
candidate is timed on both:

* early: ``probe_turns`` turns from the given particles;
* late: ``probe_turns`` turns of a repacked launch from the state after a
  pilot run of ``pilot_turns`` turns (the default shape), the lossy part of
  the launch;

and scored by the time the whole launch would take, early per-turn time x
pilot turns + late per-turn time x the remaining turns. The search is
staged: particles per lane and unroll at 256 lanes, then the workgroup size
for the best of those. Results are cached in memory and in a JSON file keyed
by device, instance count, particle count (power-of-two bucket), element
count, turns and build.
"""

from __future__ import annotations

import json
import os
import threading

from . import device, engine

BLOCKS = (64, 128, 256, 512)
KERNELS = ((1, 1), (1, 2), (1, 4), (2, 1), (2, 2))  # (particles per lane, unroll)
_CACHE_FILE = os.environ.get("ATX_TUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "at_amd_tune.json"))
_mem: dict = {}
_lock = threading.Lock()


def _key(lattices, particles, strict: bool, nturns: int, repack: bool) -> str:
    ntot = sum(int(p.shape[0]) for p in particles)
    nelem = sum(lat.nelem for lat in lattices)
    arch = engine.device_info(0)[1] if engine.device_count() else "none"
    return (f"{arch}|inst={len(lattices)}|n={1 << max(0, ntot.bit_length() - 1)}"
            f"|elems={nelem}|turns={nturns}|repack={int(repack)}|strict={int(strict)}")


def _load() -> dict:
    try:
        with open(_CACHE_FILE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    try:
        os.makedirs(os.path.dirname(_CACHE_FILE), exist_ok=True)
        tmp = _CACHE_FILE + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1)
        os.replace(tmp, _CACHE_FILE)
    except OSError:
        pass


def _timed(run, stream, repeats: int = 2) -> float:
    """Best of ``repeats`` device times [ms] of run() on ``stream``."""
    e0, e1 = device.Event(), device.Event()
    best = None
    for _ in range(repeats):
        stream.synchronize()
        e0.record(stream)
        run()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1)
        best = ms if best is None else min(best, ms)
    return best


def tune_batch(lattices, particles, *, energy: float, rest_energy: float, nturns: int = 1024,
               probe_turns: int = 3, pilot_turns: int | None = None, strict: bool = False,
               use_cache: bool = True, repack: bool = True, blocks=BLOCKS,
               kernels=KERNELS) -> dict:
    """Pick the launch shape of a batched launch of ``nturns`` turns.

    lattices: DeviceLattice per instance; particles: FP64 device buffers
    (N_i, 6) (left untouched: every probe tracks copies).
    Returns {"block_size": b, "ppt": p, "unroll": u, "ms": {shape: score},
    "early_ms": {...}, "late_ms": {...}, "pilot_turns": P}; scores are the
    modelled launch times in ms."""
    from .batch import track_batch

    key = _key(lattices, particles, strict, nturns, repack)
    with _lock:
        if use_cache and key in _mem:
            return _mem[key]
        if use_cache:
            disk = _load()
            if key in disk:
                _mem[key] = disk[key]
                return disk[key]
    if pilot_turns is None:
        pilot_turns = max(probe_turns, min(64, nturns // 8))
    pilot_turns = max(0, min(pilot_turns, nturns - 1))
    stream = device.Stream()
    common = dict(energy=energy, rest_energy=rest_energy, strict=strict, stream=stream)
    # the state after the pilot turns, with its losses (NaN rows)
    late = [p.clone(stream) for p in particles]
    if pilot_turns > 0:
        track_batch(lattices, late, pilot_turns, losses=False, repack=repack, **common)
    stream.synchronize()

    early_ms, late_ms, score = {}, {}, {}

    def measure(block, ppt, unroll):
        name = f"b{block}p{ppt}u{unroll}"
        if name in score:
            return score[name]
        shape = dict(block_size=block, ppt=ppt, unroll=unroll)

        def early():
            scratch = [p.clone(stream) for p in particles]
            track_batch(lattices, scratch, probe_turns, losses=False, **shape, **common)

        def lossy():
            scratch = [p.clone(stream) for p in late]
            track_batch(lattices, scratch, probe_turns, turn=pilot_turns, losses=False,
                        repack=repack, **shape, **common)

        early_ms[name] = _timed(early, stream)
        late_ms[name] = _timed(lossy, stream) if pilot_turns > 0 else early_ms[name]
        score[name] = (early_ms[name] * pilot_turns
                       + late_ms[name] * (nturns - pilot_turns)) / probe_turns
        return score[name]

    # particles per lane and unroll at 256 lanes, then the workgroup size
    best_k = min(kernels, key=lambda k: measure(256, *k))
    best_b = min(blocks, key=lambda b: measure(b, *best_k))
    choice = {"block_size": int(best_b), "ppt": int(best_k[0]), "unroll": int(best_k[1]),
              "ms": score, "early_ms": early_ms, "late_ms": late_ms,
              "pilot_turns": int(pilot_turns), "probe_turns": int(probe_turns)}
    with _lock:
        _mem[key] = choice
        if use_cache:
            d = _load()
            d[key] = choice
            _store(d)
    return choice
