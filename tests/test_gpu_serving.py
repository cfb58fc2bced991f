"""GPU: continuous batching (at_amd.serving).

A request served in the mix (advanced a block of turns per batched launch,
admitted while others run, loss map continued across launches) must give
exactly what the same request gives alone in one atx_track launch over all
its turns: coordinates, loss flags, loss turns, loss elements and loss
coordinates, bit for bit. Per-item turn counts of atx_track_batch and the
continued loss map are checked against single launches too."""

import numpy as np
import pytest

from at_amd import device, engine
from at_amd.device import DeviceArray
from at_amd.batch import LossMap, track_batch
from at_amd.serving import Request, Server, serve_one_by_one
from at_amd.workloads import da_polar_grid, error_realisation, gaussian_beam, large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _instances(n=3, ncells=2):
    base = large_ring(ncells=ncells, cavity=True)
    return [error_realisation(base, 500 + i) for i in range(n)]


def _alone(dl, q, energy, rest, strict=False):
    r = np.asfortranarray(q.particles.copy())
    rout, lm = dl.track(r, q.nturns, None, energy=energy, rest_energy=rest, losses=True,
                        strict=strict)
    return r, lm


def _trace(nreq=10, seed=2):
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(nreq):
        t += float(rng.exponential(0.01))
        n = int(rng.integers(50, 3000))
        nt = int(rng.choice([5, 17, 40]))
        parts = da_polar_grid(n, xmax=12e-3, ymax=5e-3) if i % 2 else gaussian_beam(n, seed=i, scale=20.0)
        out.append(Request(i, int(rng.integers(0, 3)), parts, nt, arrival=t,
                           kind="da" if i % 2 else "track"))
    return out


@pytest.mark.parametrize("repack", [True, False], ids=["repack", "plain"])
@pytest.mark.parametrize("realtime", [False, True], ids=["queued", "arrivals"])
def test_served_results_equal_alone(realtime, repack):
    lats = _instances()
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    dls = [engine.DeviceLattice(l) for l in lats]
    trace = _trace()
    srv = Server(dls, energy=energy, rest_energy=rest, turn_block=8, max_active=4,
                 hbm_budget=6000 * 105, repack=repack)
    res, st = srv.run(trace, realtime=realtime)
    print(st, flush=True)
    assert st["launches"] > 1
    nlost = 0
    for q, r in zip(trace, res):
        r_alone, lm = _alone(dls[q.instance], q, energy, rest)
        np.testing.assert_array_equal(r.particles, r_alone)
        np.testing.assert_array_equal(r.lost["islost"], lm["islost"])
        np.testing.assert_array_equal(r.lost["turn"], lm["turn"])
        np.testing.assert_array_equal(r.lost["elem"], lm["elem"])
        np.testing.assert_array_equal(r.lost["coord"], lm["coord"])
        nlost += int(lm["islost"].sum())
        assert r.done >= r.admitted - 1e-3
    assert nlost > 0  # the trace exercises the continued loss map


def test_one_by_one_baseline_equals_served():
    lats = _instances()
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    dls = [engine.DeviceLattice(l) for l in lats]
    trace = _trace(6, seed=4)
    res_b, _ = Server(dls, energy=energy, rest_energy=rest, turn_block=16).run(trace, realtime=False)
    res_1, _ = serve_one_by_one(dls, trace, energy=energy, rest_energy=rest, realtime=False)
    for a, b in zip(res_b, res_1):
        np.testing.assert_array_equal(a.particles, b.particles)
        np.testing.assert_array_equal(a.lost["turn"], b.lost["turn"])
        np.testing.assert_array_equal(a.lost["elem"], b.lost["elem"])


def test_strict_served_request_equals_oracle():
    # the strict build and the oracle agree bit for bit, RF cavity included
    lat = error_realisation(large_ring(ncells=2, cavity=True), 77)
    energy, rest = lat.energy, lat.particle.rest_energy
    dl = engine.DeviceLattice(lat)
    q = Request(0, 0, da_polar_grid(500, xmax=12e-3, ymax=5e-3), 9)
    (res,), _ = Server([dl], energy=energy, rest_energy=rest, turn_block=4,
                       strict=True).run([q], realtime=False)
    ref = np.asfortranarray(q.particles.copy())
    _, lmo = po.atpass(lat, ref, 9, losses=True)
    assert lmo["islost"].any()
    np.testing.assert_array_equal(res.lost["islost"], lmo["islost"])
    np.testing.assert_array_equal(res.lost["turn"], lmo["turn"])
    np.testing.assert_array_equal(res.lost["elem"], lmo["elem"])
    np.testing.assert_array_equal(res.lost["coord"], lmo["coord"])
    np.testing.assert_array_equal(res.particles, ref)


def test_batch_item_turns_and_continued_losses():

    lats = _instances(2)
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    dls = [engine.DeviceLattice(l) for l in lats]
    parts = [da_polar_grid(700, xmax=12e-3, ymax=5e-3), da_polar_grid(333, xmax=12e-3, ymax=5e-3)]
    tens = [DeviceArray.from_host(p.T) for p in parts]
    lms = [LossMap.empty(p.shape[1]) for p in parts]
    # instance 0: turns 0-2 then 3-9; instance 1: turns 0-6 then nothing
    track_batch(dls, tens, 0, energy=energy, rest_energy=rest, loss_maps=lms,
                item_turns=[(3, 0), (7, 0)], continue_losses=True)
    track_batch(dls, tens, 0, energy=energy, rest_energy=rest, loss_maps=lms,
                item_turns=[(7, 3), (0, 7)], continue_losses=True)
    device.synchronize()
    for i, nt in ((0, 10), (1, 7)):
        r = np.asfortranarray(parts[i].copy())
        _, lm = dls[i].track(r, nt, None, energy=energy, rest_energy=rest, losses=True)
        np.testing.assert_array_equal(tens[i].to_host().T, r)
        np.testing.assert_array_equal(lms[i].flag.to_host().astype(bool), lm["islost"])
        np.testing.assert_array_equal(lms[i].turn.to_host(), lm["turn"])
        np.testing.assert_array_equal(lms[i].elem.to_host(), lm["elem"])
        np.testing.assert_array_equal(lms[i].coord.to_host().T, lm["coord"])
        assert lm["islost"].any()


def test_budget_too_small_is_an_error():
    lats = _instances(1)
    dl = engine.DeviceLattice(lats[0])
    srv = Server([dl], energy=lats[0].energy, rest_energy=lats[0].particle.rest_energy,
                 hbm_budget=1000)
    with pytest.raises(ValueError, match="HBM budget"):
        srv.run([Request(0, 0, gaussian_beam(100), 2)], realtime=False)


def test_online_submit_from_threads_equals_alone():
    """start/submit/stop: requests submitted while the server runs (from two
    client threads, at different times) give the same results as alone."""
    import threading
    import time

    lats = _instances()
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    dls = [engine.DeviceLattice(l) for l in lats]
    trace = _trace(8, seed=6)
    srv = Server(dls, energy=energy, rest_energy=rest, turn_block=8, max_active=3,
                 hbm_budget=5000 * 105)
    srv.start()
    futs = {}

    def client(reqs):
        for q in reqs:
            futs[q.rid] = srv.submit(q)
            time.sleep(0.01)

    threads = [threading.Thread(target=client, args=(trace[k::2],)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    results = {rid: f.result(timeout=120) for rid, f in futs.items()}
    # a request over the HBM budget is refused at submit; the server goes on
    with pytest.raises(ValueError, match="over the HBM budget"):
        srv.submit(Request(999, 0, gaussian_beam(5001, seed=1), 4))
    q0 = trace[0]
    again = srv.submit(Request(1000, q0.instance, q0.particles, q0.nturns)).result(timeout=120)
    np.testing.assert_array_equal(again.particles, results[q0.rid].particles)
    srv.stop()
    with pytest.raises(RuntimeError):
        srv.submit(trace[0])
    for q in trace:
        res = results[q.rid]
        r_alone, lm = _alone(dls[q.instance], q, energy, rest)
        np.testing.assert_array_equal(res.particles, r_alone)
        np.testing.assert_array_equal(res.lost["turn"], lm["turn"])
        np.testing.assert_array_equal(res.lost["elem"], lm["elem"])
        assert res.done >= res.arrival


def test_deadline_and_cancel_end_requests_early():
    """A request whose deadline passes stops after the turns already
    launched and returns its state there, equal to tracking alone for that
    many turns; a queued request past its deadline is dropped untouched; a
    cancelled request of a live server ends the same way."""
    import time

    lats = _instances(1)
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    dl = engine.DeviceLattice(lats[0])
    long_q = Request(0, 0, gaussian_beam(20000, seed=3, scale=5.0), 100000, deadline=0.5)
    late_q = Request(1, 0, gaussian_beam(100, seed=4), 10, arrival=0.0, deadline=0.0)
    srv = Server([dl], energy=energy, rest_energy=rest, turn_block=16, max_active=1)
    (r0, r1), _ = srv.run([long_q, late_q], realtime=False)
    assert r0.status == "timeout" and 0 < r0.turns_done < long_q.nturns
    assert r0.turns_done % 16 == 0
    r = np.asfortranarray(long_q.particles.copy())
    _, lm = dl.track(r, r0.turns_done, None, energy=energy, rest_energy=rest, losses=True)
    np.testing.assert_array_equal(r0.particles, r)
    np.testing.assert_array_equal(r0.lost["turn"], lm["turn"])
    assert r1.status == "timeout" and r1.turns_done == 0
    np.testing.assert_array_equal(r1.particles, late_q.particles)

    srv2 = Server([dl], energy=energy, rest_energy=rest, turn_block=16)
    srv2.start()
    fut = srv2.submit(Request(7, 0, gaussian_beam(20000, seed=5, scale=5.0), 100000))
    time.sleep(0.5)
    srv2.cancel(7)
    res = fut.result(timeout=60)
    srv2.stop()
    assert res.status == "cancelled" and 0 < res.turns_done < 100000
    r = np.asfortranarray(gaussian_beam(20000, seed=5, scale=5.0))
    dl.track(r, res.turns_done, None, energy=energy, rest_energy=rest)
    np.testing.assert_array_equal(res.particles, r)
