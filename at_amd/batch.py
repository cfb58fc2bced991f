"""Batched tracking of many lattice instances in one persistent launch.

Each instance (an error realisation of a ring, say) has its own compiled
lattice and its own device-resident particles; all instances advance the
same number of turns in a single kernel launch (`atx_track_batch`), with
workgroups assigned to instances in proportion to their particle counts.
Particles are device buffers of shape (N_i, 6) float64 (the memory layout
of pyat's (6, N) Fortran arrays): ``device.DeviceArray``, or any object
with that shape and dtype that exposes its device address as ``ptr`` or
``data_ptr()``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import device, engine


@dataclass
class LossMap:
    """Device-resident loss map of one instance (at.c:574-594 layout)."""

    flag: device.DeviceArray   # uint8 (N,)
    turn: device.DeviceArray   # uint32 (N,)
    elem: device.DeviceArray   # uint32 (N,)
    coord: device.DeviceArray  # float64 (N, 6)

    @classmethod
    def empty(cls, n: int):
        return cls(device.DeviceArray.zeros(n, np.uint8), device.DeviceArray.zeros(n, np.uint32),
                   device.DeviceArray.zeros(n, np.uint32),
                   device.DeviceArray.zeros((n, 6), np.float64))

    def ptrs(self):
        return (self.flag.ptr, self.turn.ptr, self.elem.ptr, self.coord.ptr)

    def to_host(self) -> dict:
        """The map as lattice_track returns it: islost, turn, elem, coord (6, N)."""
        return {"islost": self.flag.to_host().astype(bool), "turn": self.turn.to_host(),
                "elem": self.elem.to_host(), "coord": self.coord.to_host().T.copy()}


def device_ptr(p) -> int:
    """Device address of a particle buffer."""
    ptr = getattr(p, "ptr", None)
    return int(ptr) if ptr is not None else int(p.data_ptr())


def _check_particles(p):
    shape = tuple(p.shape)
    dtype = np.dtype(str(p.dtype).replace("torch.", ""))
    if dtype != np.float64 or len(shape) != 2 or shape[1] != 6:
        raise ValueError("particles must be float64 device buffers of shape (N, 6)")
    if hasattr(p, "is_contiguous") and not p.is_contiguous():
        raise ValueError("particles must be contiguous")
    if hasattr(p, "is_cuda") and not p.is_cuda:
        raise ValueError("particles must be on the GPU")


def stream_handle(stream) -> int | None:
    """hipStream_t of a device.Stream, of an object with ``cuda_stream``,
    or a raw handle (None: the null stream)."""
    if stream is None:
        return None
    for attr in ("handle", "cuda_stream"):
        h = getattr(stream, attr, None)
        if h is not None:
            return int(h)
    return int(stream)


def track_batch(lattices, particles, nturns: int, *, turn: int = 0, energy: float,
                rest_energy: float, losses: bool = True, loss_maps=None, block_size: int = 0,
                ppt: int = 0, unroll: int = 0, strict: bool = False, stream=None,
                item_turns=None, continue_losses: bool = False, repack: bool = False):
    """Track every (lattice, particles) pair for ``nturns`` in one launch,
    in place. ``item_turns``, if given, holds one (nturns, turn) pair per
    instance, which then replaces the common ``nturns``/``turn``. With
    ``continue_losses`` the given loss maps carry over from an earlier call on
    the same particles. With ``repack`` the launch runs in turn blocks of 1,
    1, 2, 4, ... turns and packs each instance's surviving particles into its
    leading workgroups between blocks (one graph; identical results).
    ``block_size``, ``ppt`` (particles per lane) and ``unroll`` (slices per
    loop iteration) set the launch shape (0: the engine's default); results do
    not depend on it. Asynchronous on ``stream``. Returns the list of LossMap
    when ``losses``."""
    if len(lattices) != len(particles):
        raise ValueError("one particle buffer per lattice")
    if item_turns is not None and len(item_turns) != len(lattices):
        raise ValueError("one (nturns, turn) pair per lattice")
    for p in particles:
        _check_particles(p)
    if losses and loss_maps is None:
        loss_maps = [LossMap.empty(int(p.shape[0])) for p in particles]
    items = []
    for i, (lat, p) in enumerate(zip(lattices, particles)):
        it = {"lattice": lat, "r": device_ptr(p), "n": int(p.shape[0])}
        if item_turns is not None:
            it["nturns"], it["turn"] = (int(v) for v in item_turns[i])
            if it["nturns"] < 0:
                raise ValueError("item turn counts must be >= 0")
        if losses:
            it["lost"] = loss_maps[i].ptrs()
        items.append(it)
    engine.track_batch(items, nturns, turn=turn, energy=energy, rest_energy=rest_energy,
                       losses=losses, block_size=block_size, ppt=ppt, unroll=unroll,
                       stream=stream_handle(stream),
                       strict=strict, continue_losses=continue_losses, repack=repack)
    return loss_maps if losses else None
