"""GPU: launches made at the same time from several host threads, each on its
own stream, give the results of the same launches made one after the other.
The launch tables of every call go through shared pinned staging buffers,
reused once the copy out of them has completed; this checks that reuse under
concurrent callers."""

import threading

import numpy as np
import pytest

from at_amd import engine
from at_amd.workloads import gaussian_beam, large_ring

pytestmark = pytest.mark.gpu


def _track_blocks(dl, r, energy, rest, stream=None):
    for t in range(0, 12, 3):
        dl.track_device(r.data_ptr(), r.shape[0], 3, turn=t, energy=energy, rest_energy=rest,
                        stream=stream)


def test_threads_on_own_streams_equal_sequential():
    import torch

    ring = large_ring(ncells=2, cavity=True)
    energy, rest = ring.energy, ring.particle.rest_energy
    dev = torch.device("cuda:0")
    dls = [engine.DeviceLattice(ring) for _ in range(3)]
    beams = [gaussian_beam(3000 + 500 * k, seed=40 + k, scale=3.0) for k in range(3)]
    ref = []
    for dl, b in zip(dls, beams):
        r = torch.from_numpy(np.ascontiguousarray(b.T)).to(dev)
        _track_blocks(dl, r, energy, rest)
        torch.cuda.synchronize()
        ref.append(r.cpu().numpy())

    out = [None] * len(dls)
    errors = []

    def work(k):
        try:
            torch.cuda.set_device(dev)
            st = torch.cuda.Stream(device=dev)
            with torch.cuda.stream(st):
                r = torch.from_numpy(np.ascontiguousarray(beams[k].T)).to(dev)
                for _ in range(3):  # several rounds, so staging buffers are reused
                    r.copy_(torch.from_numpy(np.ascontiguousarray(beams[k].T)).to(dev))
                    _track_blocks(dls[k], r, energy, rest, stream=st.cuda_stream)
            st.synchronize()
            out[k] = r.cpu().numpy()
        except Exception as exc:  # noqa: BLE001 - reported below
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(dls))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    nlost = 0
    for k in range(len(dls)):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=f"thread {k}")
        nlost += int(np.isnan(ref[k][:, 0]).sum())
    print(f"{sum(b.shape[1] for b in beams)} particles, {nlost} lost", flush=True)


def test_batches_listing_lattices_in_opposite_orders():
    """Two threads batch the same two lattices, one as [A, B] and the other
    as [B, A], many times over: the engine locks each call's lattices in one
    global order, so neither waits for the other forever, and every launch
    gives the results of the same launch made alone."""
    import torch

    ring = large_ring(ncells=1, cavity=True)
    energy, rest = ring.energy, ring.particle.rest_energy
    dev = torch.device("cuda:0")
    dla, dlb = engine.DeviceLattice(ring), engine.DeviceLattice(ring)
    beam = gaussian_beam(256, seed=8, scale=3.0)
    host = np.ascontiguousarray(beam.T)

    def run(order, stream=None):
        rs = [torch.from_numpy(host).to(dev) for _ in order]
        engine.track_batch([{"lattice": dl, "r": r.data_ptr(), "n": r.shape[0]}
                            for dl, r in zip(order, rs)], 2, energy=energy, rest_energy=rest,
                           losses=False, stream=stream)
        return rs

    ref = run([dla, dlb])[0]
    torch.cuda.synchronize()
    ref = ref.cpu().numpy()
    errors, done = [], [0, 0]

    def work(k, order):
        try:
            torch.cuda.set_device(dev)
            st = torch.cuda.Stream(device=dev)
            with torch.cuda.stream(st):
                for _ in range(40):
                    rs = run(order, stream=st.cuda_stream)
                    st.synchronize()
                    for r in rs:
                        if not np.array_equal(r.cpu().numpy(), ref, equal_nan=True):
                            raise AssertionError(f"thread {k}: results differ")
                    done[k] += 1
        except Exception as exc:  # noqa: BLE001 - reported below
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(0, [dla, dlb]), daemon=True),
               threading.Thread(target=work, args=(1, [dlb, dla]), daemon=True)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), f"threads blocked after {done} launches"
    assert not errors, errors
    assert done == [40, 40]
