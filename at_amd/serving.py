"""Continuous-batching tracking server.

Lattice instances (error realisations of a ring) are compiled and uploaded
once and stay resident on the device. Requests (a particle set, an
instance, a number of turns) arrive over time. The server advances every
admitted request by up to ``turn_block`` turns per launch, all of them in
one ``atx_track_batch`` launch with per-item turn counts and turn numbers;
between launches it retires requests whose turns are done and admits
queued ones while the HBM budget and the slot limit allow. A request's
particles and loss map stay on the device from admission to completion, and
its loss map continues from launch to launch (ATX_CONTINUE_LOSSES), so the
result equals one ``atx_track`` call over all its turns, bit for bit.

Launches are queued ahead: the host knows from the turn counts which
requests a launch completes, so it queues the next launch (and the copies
back of finished requests, on a second stream) without waiting for the
device; at most ``depth`` launches are in flight. A request is done when
its results are in host memory; that time is read from a device event, so
latencies do not include host polling delays.

``serve_one_by_one`` is the baseline: requests in arrival order, one
``atx_track`` launch each over all turns.

No counterpart in the reference, whose tracking call is synchronous and
per lattice (at.c:340-681).
"""

from __future__ import annotations

import collections
import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import device, engine

BYTES_PER_PARTICLE = 48 + 1 + 4 + 4 + 48  # state + loss map (flag, turn, elem, coord)


@dataclass
class Request:
    rid: int
    instance: int             # index into the server's lattice instances
    particles: np.ndarray     # (6, N) float64
    nturns: int
    arrival: float = 0.0      # seconds after the start of the trace
    kind: str = "track"       # "track" or "da": both return coordinates and loss map
    deadline: float | None = None  # seconds after arrival; then the request is ended

    @property
    def npart(self) -> int:
        return int(self.particles.shape[1])


@dataclass
class Result:
    rid: int
    particles: np.ndarray     # (6, N) after nturns
    lost: dict                # islost, turn, elem, coord as in lattice_track's loss map
    arrival: float
    admitted: float
    done: float
    status: str = "done"      # "done", "timeout" (deadline passed) or "cancelled"
    turns_done: int = -1      # turns tracked; the request's nturns when done

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Pool:
    """Device and pinned host buffers kept for reuse, by power-of-two size
    class: freeing device memory waits for the device, and pinning host
    pages is slow, so a serving loop that allocated and freed per request
    would stall its own queued launches."""

    def __init__(self):
        self.free = collections.defaultdict(list)

    def take(self, nbytes: int):
        size = 1 << max(12, (max(nbytes, 1) - 1).bit_length())
        if self.free[size]:
            return size, self.free[size].pop()
        return size, (device.HostArray(size, np.uint8), device.DeviceArray(size, np.uint8))

    def give(self, size: int, bufs) -> None:
        self.free[size].append(bufs)


class _Block:
    """One request's device state in one allocation, mirrored by one pinned
    host buffer: particles (N, 6) f64, loss coordinates (N, 6) f64, loss
    turn and element (N,) u32, loss flag (N,) u8 -- 105 bytes per particle,
    every part at an offset aligned to its type. One copy in (particles and
    a zeroed loss map), one copy out."""

    def __init__(self, n: int, pool: _Pool):
        self.n, self.pool = n, pool
        self.nbytes = BYTES_PER_PARTICLE * n
        self.off = {"r": 0, "coord": 48 * n, "turn": 96 * n, "elem": 100 * n, "flag": 104 * n}
        self.size, (self.host, self.dev) = pool.take(self.nbytes)

    def upload(self, stream) -> None:
        device.check(device.lib().atx_memcpy(self.dev.ptr, self.host.ptr, self.nbytes,
                                             stream.handle))

    def download(self, stream) -> None:
        device.check(device.lib().atx_memcpy(self.host.ptr, self.dev.ptr, self.nbytes,
                                             stream.handle))

    def ptr(self, part: str) -> int:
        return self.dev.ptr + self.off[part]

    def view(self, part: str) -> np.ndarray:
        n, h, o = self.n, self.host.numpy, self.off[part]
        if part in ("r", "coord"):
            return h[o:o + 48 * n].view(np.float64).reshape(n, 6)
        if part in ("turn", "elem"):
            return h[o:o + 4 * n].view(np.uint32)
        return h[o:o + n]

    def stage(self, particles: np.ndarray) -> None:
        """Host side of the copy in: the particles and a zero loss map."""
        self.host.numpy[48 * self.n:self.nbytes] = 0
        self.view("r")[:] = np.asarray(particles).T

    def result(self) -> tuple:
        """(particles (6, N), loss map) from the host side after the copy out."""
        return (self.view("r").T.copy(),
                {"islost": self.view("flag").astype(bool), "turn": self.view("turn").copy(),
                 "elem": self.view("elem").copy(), "coord": self.view("coord").T.copy()})

    def close(self) -> None:
        """Back to the pool; only once the copies and launches using it
        have completed."""
        if self.pool is not None:
            self.pool.give(self.size, (self.host, self.dev))
            self.pool = None


@dataclass
class _Slot:
    req: Request
    mem: _Block
    done_turns: int = 0
    admitted: float = 0.0
    status: str = "done"


class Server:
    """Continuous batching over device-resident lattice instances.

    lattices: sequence of element sequences (one per instance).
    hbm_budget: bytes of particle state and loss maps admitted at once.
    max_active: most requests in one launch.
    turn_block: turns per launch (admission granularity).
    depth: launches queued ahead of the device.
    repack: each launch runs in turn blocks with the survivors packed
    between them (ATX_REPACK; identical results)."""

    def __init__(self, lattices, *, energy: float, rest_energy: float,
                 hbm_budget: int = 64 << 30, max_active: int = 256, turn_block: int = 128,
                 block_size: int = 0, depth: int = 2, strict: bool = False,
                 repack: bool = True):
        self.dls = [l if isinstance(l, engine.DeviceLattice) else engine.DeviceLattice(l)
                    for l in lattices]
        self.energy = float(energy)
        self.rest = float(rest_energy)
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.turn_block = int(turn_block)
        self.block_size = int(block_size)
        self.depth = max(1, int(depth))
        self.strict = bool(strict)
        self.repack = bool(repack)
        dev = C.c_int(0)
        engine._check(engine.load_library().atx_get_device(C.byref(dev)), engine.load_library())
        self.device = int(dev.value)
        self.stream = device.Stream()       # copies in and launches
        self.copy_stream = device.Stream()  # device-to-host results
        self.pool = _Pool()
        for dl in self.dls:  # compile and upload every instance now, not on first use
            dl.check(self.energy, self.rest)
        self.launches = 0
        self._cancel = set()

    def need(self, req: Request) -> int:
        return req.npart * BYTES_PER_PARTICLE

    def _admit(self, req: Request, now: float) -> _Slot:
        mem = _Block(req.npart, self.pool)
        mem.stage(req.particles)
        mem.upload(self.stream)
        return _Slot(req, mem, admitted=now)

    def _launch(self, active):
        items = []
        for s in active:
            nt = min(self.turn_block, s.req.nturns - s.done_turns)
            m = s.mem
            items.append({"lattice": self.dls[s.req.instance], "r": m.ptr("r"),
                          "n": s.req.npart, "nturns": nt, "turn": s.done_turns,
                          "lost": (m.ptr("flag"), m.ptr("turn"), m.ptr("elem"),
                                   m.ptr("coord"))})
            s.done_turns += nt
        engine.track_batch(items, 0, energy=self.energy, rest_energy=self.rest, losses=True,
                           block_size=self.block_size, stream=self.stream.handle,
                           strict=self.strict, continue_losses=True, repack=self.repack)
        self.launches += 1

    def _copy_back(self, s: _Slot):
        ev = device.Event().record(self.stream)
        self.copy_stream.wait(ev)
        s.mem.download(self.copy_stream)
        done = device.Event().record(self.copy_stream)
        return ev, done

    def run(self, requests, realtime: bool = True, log=None):
        """Serve ``requests`` (arrival times relative to the call, when
        ``realtime``; otherwise all are queued at once). Returns results in
        request order and a stats dict. ``log(result)`` is called as each
        request completes."""
        pending = collections.deque(sorted(requests, key=lambda q: (q.arrival, q.rid)))

        def fetch(tnow, block):
            out = []
            while pending and (not realtime or pending[0].arrival <= tnow):
                q = pending.popleft()
                out.append((q, q.arrival if realtime else 0.0))
            if block and not out and pending:
                time.sleep(max(0.0, min(pending[0].arrival - tnow, 0.05)))
            return out, bool(pending)

        results = {}

        def deliver(res):
            results[res.rid] = res
            if log is not None:
                log(res)

        wall = self._serve(fetch, deliver)
        res = [results[q.rid] for q in requests]
        return res, _stats(res, wall, self.energy, self.launches)

    # --- online use: requests submitted from other threads ------------------

    def start(self) -> None:
        """Serve in a background thread: ``submit`` requests while it runs,
        ``stop`` to finish. Each request's arrival is the time of its
        ``submit`` call."""
        import queue
        import threading

        if getattr(self, "_thread", None) is not None:
            raise RuntimeError("server already started")
        self._inbox = queue.Queue()
        self._stopping = threading.Event()
        self._t0 = time.perf_counter()
        self._error = None
        device = self.device

        def work():
            engine.set_device(device)
            try:
                self._serve(self._fetch_online, self._deliver_online)
            except BaseException as exc:  # noqa: BLE001 - reported through the futures
                self._error = exc
                while True:
                    try:
                        _, fut = self._inbox.get_nowait()[0:2]
                    except queue.Empty:
                        break
                    fut.set_exception(exc)
                for fut in list(self._futures.values()):
                    if not fut.done():
                        fut.set_exception(exc)

        self._futures = {}
        self._thread = threading.Thread(target=work, name="atx-server", daemon=True)
        self._thread.start()

    def submit(self, req: Request):
        """Queue a request; returns a concurrent.futures.Future of its Result."""
        from concurrent.futures import Future

        if getattr(self, "_thread", None) is None or self._stopping.is_set():
            raise RuntimeError("server is not running")
        if self._error is not None:
            raise RuntimeError("server stopped on an error") from self._error
        if self.need(req) > self.budget:
            # rejected here: admitted, it could never fit and would stop the server
            raise ValueError(f"request {req.rid} needs {self.need(req)} bytes, over the HBM "
                             f"budget of {self.budget}")
        if req.rid in self._futures:
            raise ValueError(f"request id {req.rid} is already being served")
        fut = Future()
        self._futures[req.rid] = fut
        self._inbox.put((req, fut, time.perf_counter() - self._t0))
        return fut

    def cancel(self, rid: int) -> None:
        """End a request early: a queued one is dropped, an active one stops
        after the turns already launched and returns its state at that point
        (status "cancelled"). Safe from any thread."""
        self._cancel.add(int(rid))

    def stop(self, timeout: float | None = None) -> None:
        """Serve what was submitted, then end the background thread
        (TimeoutError if it is still serving after ``timeout`` seconds;
        ``stop`` may then be called again)."""
        if getattr(self, "_thread", None) is None:
            return
        self._stopping.set()
        self._thread.join(timeout)
        if self._thread.is_alive():
            raise TimeoutError("server still serving submitted requests")
        self._thread = None
        if self._error is not None:
            raise RuntimeError("server stopped on an error") from self._error

    def _ended(self, req: Request, arrival: float, tnow: float) -> str:
        if req.rid in self._cancel:
            return "cancelled"
        if req.deadline is not None and tnow - arrival >= req.deadline:
            return "timeout"
        return ""

    def _fetch_online(self, tnow, block):
        import queue

        out = []
        try:
            if block:
                item = self._inbox.get(timeout=0.05)
                out.append(item)
            while True:
                out.append(self._inbox.get_nowait())
        except queue.Empty:
            pass
        # arrival = submit time, on the server's clock (the loop's t0 is later)
        arrivals = [(req, max(0.0, t - self._loop_t0_offset)) for req, _, t in out]
        more = not (self._stopping.is_set() and self._inbox.empty())
        return arrivals, more

    def _deliver_online(self, res):
        fut = self._futures.pop(res.rid, None)
        if fut is not None:
            fut.set_result(res)

    def _serve(self, fetch, deliver) -> float:
        """The scheduling loop. fetch(tnow, block) -> (list of (request,
        arrival), more arrivals possible); block=True when the server has
        nothing to do until a request arrives. deliver(result) receives each
        completed request. Returns the wall time."""
        queue = collections.deque()
        active: list[_Slot] = []
        finishing = collections.deque()   # (slot, finish event, copy-done event)
        inflight = collections.deque()    # launch-end events
        arrival = {}
        used = 0
        device.synchronize()
        self.stream.synchronize()
        self.copy_stream.synchronize()
        t0_ev = device.Event().record(self.stream)
        t0 = time.perf_counter()
        self._loop_t0_offset = t0 - getattr(self, "_t0", t0)

        def now():
            return time.perf_counter() - t0

        more = True
        while more or queue or active or finishing:
            tnow = now()
            new, more = fetch(tnow, False)
            for q, t in new:
                arrival[q.rid] = t
                queue.append(q)
            # deadlines and cancellations: queued requests are dropped, active
            # ones end after the turns already launched
            if self._cancel or any(q.deadline is not None for q in queue) or any(
                    sl.req.deadline is not None for sl in active):
                ended = []
                for q in list(queue):
                    why = self._ended(q, arrival[q.rid], tnow)
                    if why:
                        queue.remove(q)
                        ended.append((q, why))
                for q, why in ended:
                    self._cancel.discard(q.rid)
                    deliver(Result(q.rid, np.asarray(q.particles).copy(), _empty_loss(q.npart),
                                   arrival.pop(q.rid), tnow, tnow, why, 0))
                still = []
                for sl in active:
                    why = self._ended(sl.req, arrival[sl.req.rid], tnow)
                    if why and sl.done_turns > 0:
                        sl.status = why
                        fin, cp = self._copy_back(sl)
                        finishing.append((sl, fin, cp))
                    elif why:
                        # admitted but not launched yet: its buffers are free to go
                        self.stream.synchronize()
                        sl.mem.close()
                        used -= self.need(sl.req)
                        self._cancel.discard(sl.req.rid)
                        deliver(Result(sl.req.rid, np.asarray(sl.req.particles).copy(),
                                       _empty_loss(sl.req.npart), arrival.pop(sl.req.rid),
                                       sl.admitted, tnow, why, 0))
                    else:
                        still.append(sl)
                active = still
            while (queue and len(active) < self.max_active
                   and used + self.need(queue[0]) <= self.budget):
                q = queue.popleft()
                used += self.need(q)
                active.append(self._admit(q, max(tnow, arrival[q.rid])))
            if queue and not active and used == 0 and not finishing:
                raise ValueError(f"request {queue[0].rid} needs {self.need(queue[0])} bytes, "
                                 f"over the HBM budget of {self.budget}")
            launched = False
            if active and len(inflight) < self.depth:
                self._launch(active)
                launched = True
                inflight.append(device.Event().record(self.stream))
                still = []
                for s in active:
                    if s.done_turns >= s.req.nturns:
                        fin, cp = self._copy_back(s)
                        finishing.append((s, fin, cp))
                    else:
                        still.append(s)
                active = still
            # retire finished requests whose results are on the host
            while finishing and finishing[0][2].query():
                s, _, cp = finishing.popleft()
                done = t0_ev.elapsed_ms(cp) / 1000.0  # results on the host
                used -= self.need(s.req)
                parts, lost = s.mem.result()
                s.mem.close()
                deliver(Result(s.req.rid, parts, lost, arrival.pop(s.req.rid), s.admitted, done,
                               s.status, s.done_turns))
                self._cancel.discard(s.req.rid)
            while inflight and inflight[0].query():
                inflight.popleft()
            if launched:
                continue
            # nothing launched: wait for the device or for the next arrival
            if inflight and (len(inflight) >= self.depth or not (queue or more)):
                inflight.popleft().synchronize()
            elif finishing:
                finishing[0][2].synchronize()
            elif not queue and more:
                new, more = fetch(now(), True)
                for q, t in new:
                    arrival[q.rid] = t
                    queue.append(q)
        self.stream.synchronize()
        self.copy_stream.synchronize()
        return now()


def serve_one_by_one(dls, requests, *, energy: float, rest_energy: float, realtime: bool = True,
                     strict: bool = False, log=None):
    """Baseline: requests in arrival order, one atx_track launch each over
    all of its turns, results copied back before the next one starts."""
    dls = [l if isinstance(l, engine.DeviceLattice) else engine.DeviceLattice(l) for l in dls]
    for dl in dls:
        dl.check(energy, rest_energy)
    stream = device.Stream()
    pool = _Pool()
    order = sorted(requests, key=lambda q: (q.arrival, q.rid))
    results = {}
    device.synchronize()
    t0_ev = device.Event().record(stream)
    t0 = time.perf_counter()
    for q in order:
        if realtime:
            wait = q.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        admitted = time.perf_counter() - t0
        mem = _Block(q.npart, pool)
        mem.stage(q.particles)
        mem.upload(stream)
        dls[q.instance].track_device(mem.ptr("r"), q.npart, q.nturns, energy=energy,
                                     rest_energy=rest_energy, losses=True,
                                     loss_ptrs=(mem.ptr("flag"), mem.ptr("turn"),
                                                mem.ptr("elem"), mem.ptr("coord")),
                                     stream=stream.handle, strict=strict)
        mem.download(stream)
        fin = device.Event().record(stream)  # results on the host
        stream.synchronize()
        done = t0_ev.elapsed_ms(fin) / 1000.0
        parts, lost = mem.result()
        mem.close()
        results[q.rid] = Result(q.rid, parts, lost, q.arrival if realtime else 0.0, admitted,
                                done)
        if log is not None:
            log(results[q.rid])
    stream.synchronize()
    wall = time.perf_counter() - t0
    res = [results[q.rid] for q in requests]
    return res, _stats(res, wall, energy, len(order))


def _empty_loss(n: int) -> dict:
    return {"islost": np.zeros(n, bool), "turn": np.zeros(n, np.uint32),
            "elem": np.zeros(n, np.uint32), "coord": np.zeros((6, n))}


def _stats(results, wall: float, energy: float, launches: int) -> dict:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    return {"requests": len(results), "wall_s": wall, "launches": launches,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}


def alive_passes(requests, results, nelem_of) -> int:
    """Element passes of surviving particles over all requests (the tracked
    work, the throughput numerator): a particle lost at (turn t, element e)
    contributes t*nelem + e + 1 passes."""
    total = 0
    for q, r in zip(requests, results):
        ne = nelem_of(q.instance)
        lost = r.lost["islost"]
        nt = r.turns_done if r.turns_done >= 0 else q.nturns  # ended early: turns tracked
        total += int((~lost).sum()) * nt * ne
        if lost.any():
            t = r.lost["turn"][lost].astype(np.int64)
            e = r.lost["elem"][lost].astype(np.int64)
            total += int((t * ne + e + 1).sum())
    return total
