"""GPU: single-lattice tracking in repacked turn blocks (atx_track with
ATX_REPACK; gpupass ``repack=True``, which the acceptance drivers use)
gives the results of one plain launch bit for bit: final coordinates, loss
flags, turns, elements and coordinates."""

import numpy as np
import pytest

from at_amd import device, engine, tracking
from at_amd.batch import LossMap
from at_amd.workloads import da_polar_grid, large_ring

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ring():
    return large_ring()


def _run(dl, ring, beam, nturns, repack, strict, start=0, turn=0):
    r = np.asfortranarray(beam.copy())
    rout, lm = dl.track(r, nturns, None, turn=turn, energy=ring.energy,
                        rest_energy=ring.particle.rest_energy, losses=True, start_elem=start,
                        strict=strict, repack=repack)
    return r, lm


@pytest.mark.parametrize("strict", (False, True))
def test_repacked_single_lattice_equals_plain(ring, strict):
    beam = da_polar_grid(6000)
    dl = engine.DeviceLattice(ring)
    try:
        for nturns, start, turn in ((40, 0, 0), (9, 100, 7)):
            print(f"strict={strict} nturns={nturns} start_elem={start}", flush=True)
            a, la = _run(dl, ring, beam, nturns, False, strict, start, turn)
            b, lb = _run(dl, ring, beam, nturns, True, strict, start, turn)
            assert la["islost"].sum() > 100  # the DA grid loses particles
            np.testing.assert_array_equal(a, b)
            for k in ("islost", "turn", "elem", "coord"):
                np.testing.assert_array_equal(la[k], lb[k], err_msg=k)
    finally:
        dl.close()


def test_device_pointers_and_lattice_track(ring):
    """Device-resident particles with ATX_REPACK equal a plain launch, and so
    does lattice_track with repack=True."""
    beam = da_polar_grid(3000)
    n = beam.shape[1]
    dl = engine.DeviceLattice(ring)
    try:
        outs = []
        for repack in (False, True):
            r = device.DeviceArray.from_host(beam.T)
            lm = LossMap.empty(n)
            dl.track_device(r.ptr, n, 24, energy=ring.energy,
                            rest_energy=ring.particle.rest_energy, losses=True,
                            loss_ptrs=lm.ptrs(), repack=repack)
            device.synchronize()
            outs.append((r.to_host(), lm.to_host()))
            r.close()
        np.testing.assert_array_equal(outs[0][0], outs[1][0])
        for k in ("islost", "turn", "elem", "coord"):
            np.testing.assert_array_equal(outs[0][1][k], outs[1][1][k], err_msg=k)
    finally:
        dl.close()
    lat = list(ring)
    res = []
    for repack in (True, False):
        r = np.asfortranarray(beam.copy())
        _, _, data = tracking.lattice_track(lat, r, 64, refpts=None, in_place=True,
                                            use_mp=tracking.MPMode.GPU,
                                            losses=True, energy=ring.energy,
                                            particle=ring.particle, repack=repack)
        res.append((r, data["loss_map"]))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for k in ("islost", "turn", "elem"):
        np.testing.assert_array_equal(res[0][1][k], res[1][1][k], err_msg=k)
