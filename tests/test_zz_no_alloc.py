"""GPU: atx_track_batch_noalloc, the entry point that allocates nothing,
against atx_track_batch on the same items: identical coordinates and loss
maps, bit for bit, on two calls in a row, the second with fewer items and
fewer particles (its fixed instance table keeps the first call's entries
beyond the second call's items; they must not matter)."""

import ctypes as C

import numpy as np
import pytest

from at_amd import device, engine
from at_amd.batch import LossMap
from at_amd.workloads import da_polar_grid, error_realisation, large_ring

pytestmark = pytest.mark.gpu


def _run(dls, parts, turns, noalloc, strict):
    bufs = [device.DeviceArray.from_host(p.T) for p in parts]
    lms = [LossMap.empty(p.shape[1]) for p in parts]
    items = [{"lattice": dl, "r": b.ptr, "n": b.shape[0], "nturns": nt, "turn": t0,
              "lost": lm.ptrs()} for dl, b, lm, (nt, t0) in zip(dls, bufs, lms, turns)]
    engine.track_batch(items, 0, energy=dls[0].energy, rest_energy=dls[0].rest, losses=True,
                       strict=strict, noalloc=noalloc)
    device.synchronize()
    return [b.to_host() for b in bufs], [lm.to_host() for lm in lms]


@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_noalloc_equals_batch_on_two_calls(strict):
    base = large_ring(ncells=2, cavity=True)
    lats = [error_realisation(base, 900 + i) for i in range(5)]
    dls = [engine.DeviceLattice(l) for l in lats]
    for dl in dls:
        dl.energy, dl.rest = base.energy, base.particle.rest_energy
    sizes = [1500, 700, 2100, 333, 900]
    parts = [da_polar_grid(n, xmax=12e-3, ymax=5e-3) for n in sizes]
    turns = [(9, 0), (4, 3), (12, 0), (7, 1), (1, 0)]
    # the lattices go to the device with one ordinary call
    ref = _run(dls, parts, turns, False, strict)
    print(f"strict={strict}: first call, {len(dls)} items", flush=True)
    got = _run(dls, parts, turns, True, strict)
    nlost = 0
    for i in range(len(dls)):
        np.testing.assert_array_equal(got[0][i], ref[0][i], err_msg=f"item {i}")
        for k in ref[1][i]:
            np.testing.assert_array_equal(got[1][i][k], ref[1][i][k], err_msg=f"item {i} {k}")
        nlost += int(got[1][i]["islost"].sum())
    assert nlost > 0
    # second call: fewer items, fewer particles
    sub = [1, 3]
    parts2 = [p[:, : p.shape[1] // 2].copy(order="F") for p in (parts[i] for i in sub)]
    d2 = [dls[i] for i in sub]
    t2 = [(6, 2), (3, 0)]
    print(f"strict={strict}: second call, {len(d2)} items", flush=True)
    ref2 = _run(d2, parts2, t2, False, strict)
    got2 = _run(d2, parts2, t2, True, strict)
    for i in range(len(d2)):
        np.testing.assert_array_equal(got2[0][i], ref2[0][i], err_msg=f"item {i}")
        for k in ref2[1][i]:
            np.testing.assert_array_equal(got2[1][i][k], ref2[1][i][k], err_msg=f"item {i} {k}")


def test_noalloc_limits_are_errors():
    base = large_ring(ncells=1, cavity=True)
    dl = engine.DeviceLattice(base)
    lib = engine.load_library()
    items = (engine.BatchItem * 65)()
    for it in items:
        it.lattice = dl.h
        it.num_particles = 0
    rc = lib.atx_track_batch_noalloc(items, 65, 1, 0, base.energy, base.particle.rest_energy,
                                     0, None, None, 0)
    assert rc == engine.ATX_E_VALUE and b"at most 64" in lib.atx_last_error()
    rc = lib.atx_track_batch_noalloc(items, 1, 1, 0, base.energy, base.particle.rest_energy,
                                     0, None, None, engine.ATX_REPACK)
    assert rc == engine.ATX_E_VALUE and b"ATX_REPACK" in lib.atx_last_error()
    # a lattice never sent to the device
    b = device.DeviceArray.from_host(da_polar_grid(64).T)
    items[0].r = b.ptr
    items[0].num_particles = 64
    rc = lib.atx_track_batch_noalloc(items, 1, 1, 0, base.energy, base.particle.rest_energy,
                                     0, None, None, 0)
    assert rc == engine.ATX_E_VALUE and b"not on this device" in lib.atx_last_error()
