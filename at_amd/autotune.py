"""Launch-shape autotuner for batched tracking.

The per-turn cost of the tracking kernel is the same every turn, so a few
probe turns on scratch copies of the particle buffers rank the candidate
workgroup sizes reliably. Results are cached in memory and in a JSON file
keyed by device, instance count, particle count (power-of-two bucket) and
element count.
"""

from __future__ import annotations

import json
import os
import threading

from . import device, engine

CANDIDATES = (64, 128, 256, 512)
_CACHE_FILE = os.environ.get("ATX_TUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "at_amd_tune.json"))
_mem: dict = {}
_lock = threading.Lock()


def _key(lattices, particles, strict: bool) -> str:
    ntot = sum(int(p.shape[0]) for p in particles)
    nelem = sum(lat.nelem for lat in lattices)
    arch = engine.device_info(0)[1] if engine.device_count() else "none"
    return (f"{arch}|inst={len(lattices)}|n={1 << max(0, ntot.bit_length() - 1)}"
            f"|elems={nelem}|strict={int(strict)}")


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


def tune_batch(lattices, particles, *, energy: float, rest_energy: float, probe_turns: int = 3,
               candidates=CANDIDATES, strict: bool = False, use_cache: bool = True,
               repack: bool = False) -> dict:
    """Pick the workgroup size for one batched launch.

    lattices: DeviceLattice per instance; particles: FP64 device buffers
    (N_i, 6) (left untouched: probes track clones).
    Returns {"block_size": b, "ms": {b: best probe time}}."""
    from .batch import track_batch

    key = _key(lattices, particles, strict)
    with _lock:
        if use_cache and key in _mem:
            return _mem[key]
        if use_cache:
            disk = _load()
            if key in disk:
                _mem[key] = disk[key]
                return disk[key]
    stream = device.Stream()
    e0, e1 = device.Event(), device.Event()
    times = {}
    for b in candidates:
        best = None
        for _ in range(2):
            scratch = [p.clone() for p in particles]
            device.synchronize()
            e0.record(stream)
            track_batch(lattices, scratch, probe_turns, energy=energy, rest_energy=rest_energy,
                        losses=False, block_size=b, strict=strict, stream=stream)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1)
            best = ms if best is None else min(best, ms)
        times[b] = best
    choice = {"block_size": int(min(times, key=times.get)),
              "ms": {str(k): v for k, v in times.items()}}
    with _lock:
        _mem[key] = choice
        if use_cache:
            d = _load()
            d[key] = choice
            _store(d)
    return choice
