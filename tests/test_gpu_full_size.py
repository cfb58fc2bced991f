"""GPU: the headline workload at its full size -- the large ring (2,241
elements, radiation on, cavity on) with the 262,144-particle Gaussian beam
(scale 10) in the default launch shape. The oracle cannot track the whole
beam in a test's time, so the full-size runs are checked through properties
that do not depend on the beam size:

* a particle's result does not depend on the beam it travels with: a sample
  of particles tracked alone equals the same particles inside the full
  launch, bit for bit (coordinates and loss records), in both builds --
  the lost-particle compaction, the workgroup a particle lands in and the
  launch size change nothing in its arithmetic;
* on that sample the strict build equals the oracle bit for bit, and the
  fast build has the oracle's loss records and its coordinates within
  1e-11;
* turns compose: 3 turns in one call equal 1 turn and then 2 more from
  turn 1, bit for bit, including the loss records.

The batched workload at its full size (64 orbit-corrected error
realisations of the large ring, DA grids of 1,400,590 particles in all, one
repacked launch in XCD order) is checked the same way: sampled instances
equal their own single-lattice launches bit for bit in both builds, and in
the strict build a sample of one instance equals the oracle."""

import numpy as np
import pytest

from at_amd import device, engine
from at_amd.batch import track_batch
from at_amd.workloads import batched_da_workload, gaussian_beam, large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

N = 262144
TURNS = 3


@pytest.fixture(scope="module")
def setup():
    ring = large_ring()
    beam = gaussian_beam(N, seed=0, scale=10.0)
    dl = engine.DeviceLattice(ring)
    # a sample across the whole beam: the first workgroup's particles and a
    # stride through the rest
    sample = np.unique(np.concatenate([np.arange(600), np.arange(600, N, 509)]))
    return ring, beam, dl, sample


def _track(dl, ring, r0, nturns, strict, turn=0):
    r = np.asfortranarray(r0.copy())
    _, lm = dl.track(r, nturns, None, turn=turn, energy=ring.energy,
                     rest_energy=ring.particle.rest_energy, losses=True, strict=strict)
    return r, lm


def _records(lm):
    return {k: np.asarray(lm[k]) for k in ("islost", "turn", "elem", "coord")}


@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_full_beam_equals_particles_alone_and_oracle(setup, strict):
    ring, beam, dl, sample = setup
    print(f"full size: {N} particles, {TURNS} turns, strict={strict}", flush=True)
    full, lm_full = _track(dl, ring, beam, TURNS, strict)
    # with every particle the full launch lost, so that loss records are
    # compared as well
    sample = np.union1d(sample, np.flatnonzero(lm_full["islost"]))
    part, lm_part = _track(dl, ring, beam[:, sample], TURNS, strict)
    np.testing.assert_array_equal(full[:, sample], part)
    rf, rp = _records(lm_full), _records(lm_part)
    for k in ("islost", "turn", "elem"):
        np.testing.assert_array_equal(rf[k][sample], rp[k], err_msg=k)
    np.testing.assert_array_equal(rf["coord"][:, sample], rp["coord"], err_msg="coord")
    nlost = int(rf["islost"].sum())
    print(f"full size: {nlost} of {N} lost in {TURNS} turns; sample of {sample.size}", flush=True)
    assert nlost > 0 and int(rp["islost"].sum()) == nlost
    ref = np.asfortranarray(beam[:, sample].copy())
    _, lmo = po.atpass(ring, ref, TURNS, losses=True)
    for k in ("islost", "turn", "elem"):
        np.testing.assert_array_equal(rp[k], np.asarray(lmo[k]), err_msg=k)
    if strict:
        np.testing.assert_array_equal(part, ref)
    else:
        # the fast build's bar: 1e-12 per coordinate after one turn; 1e-11
        # after these three
        surv = ~rp["islost"]
        np.testing.assert_allclose(part[:, surv], ref[:, surv], rtol=0, atol=1e-11)


@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_full_beam_turns_compose(setup, strict):
    ring, beam, dl, _ = setup
    whole, lm_whole = _track(dl, ring, beam, TURNS, strict)
    r = np.asfortranarray(beam.copy())
    e, rest = ring.energy, ring.particle.rest_energy
    _, lm1 = dl.track(r, 1, None, turn=0, energy=e, rest_energy=rest, losses=True, strict=strict)
    _, lm2 = dl.track(r, TURNS - 1, None, turn=1, energy=e, rest_energy=rest, losses=True,
                      strict=strict)
    np.testing.assert_array_equal(r, whole)
    a, b1, b2 = _records(lm_whole), _records(lm1), _records(lm2)
    first = b1["islost"].astype(bool)
    # particles lost in the first call keep that call's record; the others
    # carry the second call's
    for k in ("turn", "elem"):
        merged = np.where(first, b1[k], b2[k])
        np.testing.assert_array_equal(a[k], merged, err_msg=k)
    np.testing.assert_array_equal(a["islost"].astype(bool), first | b2["islost"].astype(bool))


@pytest.fixture(scope="module")
def realisations():
    return batched_da_workload(ninst=64, seed=7, shift=None)


@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_batched_realisations_full_size_equal_single_launches(realisations, strict):
    work = realisations
    lats = [l for l, _ in work]
    parts = [p for _, p in work]
    print(f"full-size batch: {len(lats)} instances, {sum(p.shape[1] for p in parts)} particles",
          flush=True)
    dls = [engine.DeviceLattice(l) for l in lats]
    e, rest = lats[0].energy, lats[0].particle.rest_energy
    tens = [device.DeviceArray.from_host(p.T) for p in parts]
    nturns = 6
    lms = track_batch(dls, tens, nturns, energy=e, rest_energy=rest, block_size=64, ppt=1,
                      unroll=1 if strict else 2, repack=True, strict=strict)
    device.synchronize()
    nlost = 0
    for i in (0, 17, 40, 63):
        alone = np.asfortranarray(parts[i].copy())
        _, lm1 = dls[i].track(alone, nturns, None, energy=e, rest_energy=rest, losses=True,
                              strict=strict)
        np.testing.assert_array_equal(tens[i].to_host().T, alone, err_msg=f"instance {i}")
        flag = lms[i].flag.to_host().astype(bool)
        np.testing.assert_array_equal(flag, lm1["islost"], err_msg=f"instance {i}")
        np.testing.assert_array_equal(lms[i].turn.to_host(), lm1["turn"], err_msg=f"instance {i}")
        np.testing.assert_array_equal(lms[i].elem.to_host(), lm1["elem"], err_msg=f"instance {i}")
        nlost += int(flag.sum())
        if strict and i == 0:
            # a sample of the instance against the oracle, bit for bit
            pick = np.unique(np.concatenate([np.arange(0, parts[i].shape[1], 97),
                                             np.flatnonzero(flag)]))
            ref = np.asfortranarray(parts[i][:, pick].copy())
            _, lmo = po.atpass(lats[i], ref, nturns, losses=True)
            np.testing.assert_array_equal(alone[:, pick], ref)
            for k in ("islost", "turn", "elem"):
                np.testing.assert_array_equal(lm1[k][pick], np.asarray(lmo[k]), err_msg=k)
    assert nlost > 0
