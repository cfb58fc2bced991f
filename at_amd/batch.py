"""Batched tracking of many lattice instances in one persistent launch.

Each instance (an error realisation of a ring, say) has its own compiled
lattice and its own device-resident particles; all instances advance the
same number of turns in a single kernel launch (`atx_track_batch`), with
workgroups assigned to instances in proportion to their particle counts.
Particles are torch FP64 tensors of shape (N_i, 6) on the device (the
memory layout of pyat's (6, N) Fortran arrays).
"""

from __future__ import annotations

from dataclasses import dataclass

from . import engine


@dataclass
class LossMap:
    """Device-resident loss map of one instance (at.c:574-594 layout)."""

    flag: "object"   # torch.uint8 (N,)
    turn: "object"   # torch.int32 (N,), uint32 values
    elem: "object"   # torch.int32 (N,)
    coord: "object"  # torch.float64 (N, 6)

    @classmethod
    def empty(cls, n: int, device):
        import torch

        return cls(torch.zeros(n, dtype=torch.uint8, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros(n, dtype=torch.int32, device=device),
                   torch.zeros((n, 6), dtype=torch.float64, device=device))

    def ptrs(self):
        return (self.flag.data_ptr(), self.turn.data_ptr(), self.elem.data_ptr(),
                self.coord.data_ptr())


def _check_particles(p):
    import torch

    if p.dtype != torch.float64 or p.dim() != 2 or p.shape[1] != 6 or not p.is_contiguous():
        raise ValueError("particles must be contiguous float64 tensors of shape (N, 6)")
    if not p.is_cuda:
        raise ValueError("particles must be on the GPU")


def track_batch(lattices, particles, nturns: int, *, turn: int = 0, energy: float,
                rest_energy: float, losses: bool = True, loss_maps=None, block_size: int = 0,
                strict: bool = False, stream=None, item_turns=None,
                continue_losses: bool = False, repack: bool = False):
    """Track every (lattice, particles) pair for ``nturns`` in one launch,
    in place. ``item_turns``, if given, holds one (nturns, turn) pair per
    instance, which then replaces the common ``nturns``/``turn``. With
    ``continue_losses`` the given loss maps carry over from an earlier call on
    the same particles. With ``repack`` the launch runs in turn blocks of 1,
    1, 2, 4, ... turns and packs each instance's surviving particles into its
    leading workgroups between blocks (one graph; identical results). Returns
    the list of LossMap when ``losses``."""
    import torch

    if len(lattices) != len(particles):
        raise ValueError("one particle tensor per lattice")
    if item_turns is not None and len(item_turns) != len(lattices):
        raise ValueError("one (nturns, turn) pair per lattice")
    for p in particles:
        _check_particles(p)
    if losses and loss_maps is None:
        loss_maps = [LossMap.empty(p.shape[0], p.device) for p in particles]
    items = []
    for i, (lat, p) in enumerate(zip(lattices, particles)):
        it = {"lattice": lat, "r": p.data_ptr(), "n": p.shape[0]}
        if item_turns is not None:
            it["nturns"], it["turn"] = (int(v) for v in item_turns[i])
            if it["nturns"] < 0:
                raise ValueError("item turn counts must be >= 0")
        if losses:
            it["lost"] = loss_maps[i].ptrs()
        items.append(it)
    s = stream if stream is not None else torch.cuda.current_stream()
    engine.track_batch(items, nturns, turn=turn, energy=energy, rest_energy=rest_energy,
                       losses=losses, block_size=block_size, stream=s.cuda_stream,
                       strict=strict, continue_losses=continue_losses, repack=repack)
    return loss_maps if losses else None
