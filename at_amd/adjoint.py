"""Gradients of a tracking loss with respect to magnet strengths.

``TrackingLoss`` is a torch.autograd.Function whose forward tracks the
particles through ``nturns`` turns on the engine and returns
sum over surviving particles of x^2 + px^2 + y^2 + py^2 at the end; its
backward is the hand-written adjoint kernel (atx_adjoint):

* forward keeps one checkpoint of the particle state every ``checkpoint``
  turns (8 by default);
* backward first recomputes every block forward from its checkpoint, all
  blocks together: turn j of every block is one batched launch
  (atx_track_batch, one instance per block with its own turn number), and
  the state at the start of each turn is kept; the blocks' arithmetic is
  that of the forward pass, so the states are bit-identical to it;
* it then runs the adjoint kernel over the blocks, which reconstructs every
  element's input by exact inversion and accumulates dL/dPolynomB[k] of the
  selected elements. Sequentially, block after block from the last; or,
  when the beam is too small to fill the device, in parallel: the adjoint
  is linear in lambda, so one launch takes unit seeds e_0..e_5 through every
  block but the first (rows of each block's transposed Jacobian, one 6x6
  matrix per particle) together with the last block's own sweep; lambda at
  every block boundary then follows by 6x6 products, and a second launch
  sweeps the remaining blocks together, accumulating the gradient. That is
  two launches instead of nblocks in sequence, for about 7 times the lane
  work. The parallel items are sub-blocks of g turns of the checkpoint
  blocks (their turn-start states are recorded), g chosen by a cost model
  (sweep_plan).

Memory: checkpoints (nturns/checkpoint + 1) x N x 48 B plus the turn-start
states of all blocks, (nturns + nblocks) x N x 48 B (12.6 MB for 4,096
particles and 64 turns); the parallel sweep adds (S - 1) x N x 288 B of
sub-block matrices for S sub-blocks.
"""

from __future__ import annotations

import numpy as np

from . import engine


class Problem:
    """A lattice, its differentiation parameters and the tracking setup.

    params: sequence of (element index, PolynomB index). parallel: sweep
    the blocks in parallel (True), in sequence (False) or as a cost model of
    the device's occupancy prefers (None). sub_block: turns per item of the
    parallel sweep (None: chosen with the sweep by the cost model; 0:
    sequential)."""

    def __init__(self, lattice, params, *, energy=None, rest_energy=None, checkpoint: int = 8,
                 parallel: bool | None = None, sub_block: int | None = None):
        self.lattice = lattice
        self.params = [(int(e), int(k)) for e, k in params]
        self.energy = float(lattice.energy if energy is None else energy)
        part = getattr(lattice, "particle", None)
        self.rest = float(getattr(part, "rest_energy", 0.0) if rest_energy is None else rest_energy)
        self.checkpoint = int(checkpoint)
        self.parallel = parallel
        self.sub_block = None if sub_block is None else int(sub_block)
        self.dl = engine.DeviceLattice(lattice)
        n = len(lattice)
        self.slots = np.full(n, -1, np.int32)
        self.coeffs = np.zeros(n, np.int32)
        for i, (e, k) in enumerate(self.params):
            if self.slots[e] >= 0:
                raise ValueError(f"element {e} appears twice in params")
            self.slots[e] = i
            self.coeffs[e] = k
        self._pb = {e: np.array(lattice[e].PolynomB, dtype=np.float64) for e, _ in self.params}
        self._current = self.initial_strengths()  # strengths on the device lattice

    def initial_strengths(self) -> np.ndarray:
        return np.array([self._pb[e][k] for e, k in self.params])

    def set_strengths(self, values) -> None:
        """Put these strengths on the device lattice (recompiled on its next
        launch); nothing happens when they are already there."""
        vals = np.array(values, dtype=np.float64)
        if vals.shape == self._current.shape and np.array_equal(vals, self._current):
            return
        self._current = vals
        touched = {}
        for (e, k), v in zip(self.params, vals):
            pb = touched.setdefault(e, self._pb[e].copy())
            pb[k] = v
        for e, pb in touched.items():
            self.dl.set_array_at(e, "PolynomB", pb)


def _sweep_cost(lanes: int, simds: int) -> float:
    """Time per turn of one adjoint launch over `lanes` particles, in units
    of a launch with at most one wave per SIMD: below that the dependent FP64
    chain of each lane sets the time; above it the time grows as
    0.536 + 0.533 w with w waves per SIMD (fitted to launches over the large
    ring of 4,096-particle items: 1.57 at w = 1.9, 1.97 at w = 2.7, 6.77 at
    w = 11.7; it orders 1-, 2- and 4-turn sub-blocks as measured)."""
    return max(1.0, 0.536 + 0.533 * lanes / (64.0 * simds))


def sweep_plan(npart: int, block_turns, simds: int) -> int:
    """Sub-block length of the parallel sweep, in turns, or 0 for the
    sequential one. The parallel sweep splits the checkpoint blocks (turn
    counts block_turns) into S sub-blocks of g turns (their turn-start states
    are all recorded) and makes two launches of g turns, one of
    6 (S - 1) + 1 items and one of S - 1; the sequential sweep walks every
    turn with npart lanes. Shorter sub-blocks mean more lanes and shorter
    launches; the cost model picks g."""
    turns = int(sum(block_turns))
    if len(block_turns) < 1 or turns < 2:
        return 0
    seq = turns * _sweep_cost(npart, simds)
    best, best_cost = 0, 0.8 * seq
    g = 1
    while g <= max(block_turns):
        nsub = sum(-(-nt // g) for nt in block_turns)
        if nsub >= 2:
            cost = g * (_sweep_cost((6 * (nsub - 1) + 1) * npart, simds)
                        + _sweep_cost((nsub - 1) * npart, simds))
            if cost < best_cost:
                best, best_cost = g, cost
        g *= 2
    return best


def use_parallel_sweep(npart: int, nblocks: int, simds: int, checkpoint: int = 8) -> bool:
    """Whether the cost model prefers the parallel sweep for nblocks
    checkpoint blocks of `checkpoint` turns."""
    return sweep_plan(npart, [checkpoint] * nblocks, simds) > 0


def _device_simds(dev) -> int:
    import torch

    return 4 * torch.cuda.get_device_properties(dev).multi_processor_count


def _sweep_sequential(problem, blocks, states, lam, grad, n, stream):
    npar = len(problem.params)
    for k in reversed(range(len(blocks))):
        t0, nt, _ = blocks[k]
        problem.dl.adjoint(states[k].data_ptr(), nt, t0, n, problem.energy, problem.rest,
                           lam.data_ptr(), problem.slots, problem.coeffs, grad.data_ptr(), npar,
                           stream=stream)


def _sub_blocks(blocks, states, g: int):
    """Sub-blocks of at most g turns of the checkpoint blocks: (turn-start
    states of the sub-block and its end state, a contiguous view; turn
    number of its first turn; its turn count), in turn order."""
    subs = []
    for k, (t0, nt, _) in enumerate(blocks):
        for j in range(0, nt, g):
            m = min(g, nt - j)
            subs.append((t0 + j, m, states[k][j:j + m + 1]))
    return subs


def _sweep_parallel(problem, blocks, states, lam, grad, n, stream, g: int = 0):
    """Sub-blocks of g turns (whole blocks for g = 0) in parallel: unit
    seeds through sub-blocks 1..S-1 (with the last one's own sweep in the
    same launch), lambda at the boundaries by 6x6 products, then sub-blocks
    0..S-2 in one launch."""
    import torch

    if g > 0:
        blocks = _sub_blocks(blocks, states, g)
        states = [st for _, _, st in blocks]
    nb = len(blocks)
    npar = len(problem.params)
    # seeds[k - 1, i]: adjoint of the end of block k seeded with e_i; after
    # the launch, row i of block k's transposed Jacobian for every particle
    seeds = torch.zeros((nb - 1, 6, n, 6), dtype=torch.float64, device=lam.device)
    for i in range(6):
        seeds[:, i, :, i] = 1.0
    items = []
    for k in range(1, nb):
        t0, nt, _ = blocks[k]
        for i in range(6):
            items.append((states[k].data_ptr(), seeds[k - 1, i].data_ptr(), nt, t0, n, False))
    t0, nt, _ = blocks[-1]
    items.append((states[-1].data_ptr(), lam.data_ptr(), nt, t0, n, True))
    problem.dl.adjoint_batch(items, problem.energy, problem.rest, problem.slots, problem.coeffs,
                             grad.data_ptr(), npar, stream=stream)
    # ends[k]: lambda at the end of block k (= at the start of block k + 1)
    ends = {nb - 2: lam}
    for k in range(nb - 2, 0, -1):
        # rows[p, i]: lambda at the start of block k for lambda = e_i at its end
        rows = seeds[k - 1].permute(1, 0, 2)
        ends[k - 1] = (ends[k][:, :, None] * rows).sum(dim=1).contiguous()
    items = [(states[k].data_ptr(), ends[k].data_ptr(), blocks[k][1], blocks[k][0], n, True)
             for k in range(nb - 1)]
    problem.dl.adjoint_batch(items, problem.energy, problem.rest, problem.slots, problem.coeffs,
                             grad.data_ptr(), npar, stream=stream)


def _loss_and_seed(r_end):
    """Loss of the end state and its adjoint (zero for lost particles)."""
    import torch

    alive = torch.isfinite(r_end).all(dim=1)
    q = r_end[:, :4]
    loss = torch.where(alive[:, None], q * q, torch.zeros_like(q)).sum()
    lam = torch.zeros_like(r_end)
    lam[:, :4] = torch.where(alive[:, None], 2.0 * q, torch.zeros_like(q))
    return loss, lam


class TrackingLoss:
    """Differentiable L(strengths) = sum_alive (x^2 + px^2 + y^2 + py^2)."""

    @staticmethod
    def apply(problem: Problem, strengths, particles, nturns: int):
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta):
                strengths_np = theta.detach().cpu().numpy().copy()
                problem.set_strengths(strengths_np)
                r = particles.detach().clone()
                n = r.shape[0]
                cp = problem.checkpoint
                ckpts = []
                stream = torch.cuda.current_stream().cuda_stream
                t = 0
                # a beam of at most one wave per SIMD runs in 64-thread
                # workgroups, its waves on four times as many CUs (the large
                # ring's 8-turn block of 4,096 particles: 19.0 ms, against
                # 20.3 ms in 256-thread workgroups)
                small = n <= 64 * _device_simds(r.device)
                while t < nturns:
                    nt = min(cp, nturns - t)
                    ckpts.append((t, nt, r.clone()))
                    engine.track_batch([{"lattice": problem.dl, "r": r.data_ptr(), "n": n,
                                         "nturns": nt, "turn": t}], 0, energy=problem.energy,
                                       rest_energy=problem.rest, losses=False,
                                       block_size=64 if small else 0, stream=stream)
                    t += nt
                loss, lam = _loss_and_seed(r)
                ctx.ckpts = ckpts
                ctx.strengths = strengths_np
                ctx.lam = lam
                ctx.theta_device = theta.device
                ctx.final = r
                return loss

            @staticmethod
            def backward(ctx, grad_out):
                n = particles.shape[0]
                npar = len(problem.params)
                grad = torch.zeros(npar, dtype=torch.float64, device=particles.device)
                lam = (ctx.lam * grad_out).contiguous()
                stream = torch.cuda.current_stream().cuda_stream
                blocks = ctx.ckpts
                if not blocks:
                    return grad.to(device=ctx.theta_device, dtype=torch.float64)
                # the lattice is shared: another forward may have set other
                # strengths since this one; recompute and sweep with this one's
                problem.set_strengths(ctx.strengths)
                states = [torch.empty((nt + 1, n, 6), dtype=torch.float64,
                                      device=particles.device) for _, nt, _ in blocks]
                cur = [ck.clone() for _, _, ck in blocks]
                for j in range(max(nt for _, nt, _ in blocks)):
                    items = []
                    for k, (t0, nt, _) in enumerate(blocks):
                        if j < nt:
                            states[k][j].copy_(cur[k])
                            items.append({"lattice": problem.dl, "r": cur[k].data_ptr(), "n": n,
                                          "nturns": 1, "turn": t0 + j})
                    engine.track_batch(items, 0, energy=problem.energy,
                                       rest_energy=problem.rest, losses=False, stream=stream)
                for k, (_, nt, _) in enumerate(blocks):
                    states[k][nt].copy_(cur[k])
                g = problem.sub_block
                if g is None:
                    g = sweep_plan(n, [nt for _, nt, _ in blocks], _device_simds(particles.device))
                    if problem.parallel is True and g == 0:
                        g = problem.checkpoint
                    elif problem.parallel is False:
                        g = 0
                if g > 0 and sum(-(-nt // g) for _, nt, _ in blocks) > 1:
                    _sweep_parallel(problem, blocks, states, lam, grad, n, stream, g)
                else:
                    _sweep_sequential(problem, blocks, states, lam, grad, n, stream)
                return grad.to(device=ctx.theta_device, dtype=torch.float64)

        return _F.apply(strengths)


def loss_and_grad(problem: Problem, particles_np: np.ndarray, nturns: int, strengths=None):
    """Convenience: (loss, dL/dstrengths) as numpy for a (6, N) beam."""
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    parts = torch.from_numpy(np.ascontiguousarray(np.asarray(particles_np).T)).to(dev)
    theta0 = problem.initial_strengths() if strengths is None else np.asarray(strengths)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = TrackingLoss.apply(problem, theta, parts, nturns)
    loss.backward()
    return float(loss.detach().cpu()), theta.grad.detach().cpu().numpy()
