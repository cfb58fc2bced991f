"""GPU: the launch shape (workgroup size, particles per lane, slices per
loop iteration; include/atx.h atx_launch_shape) changes how the kernel is
scheduled, never what it computes.

Every shape gives bit for bit the result of the default shape, in the
strict and the fast build, with losses (compaction of survivors between
turns), with refpts (no compaction, lost particles keep filling their
copies), with the loss map off (the isfinite(x) quirk of setlost), and with
the inputs at.c treats specially (NaN, Inf, |r| > 1). The strict build
equals the oracle for every shape."""

import numpy as np
import pytest

import zoo
from at_amd import device, engine
from at_amd.batch import track_batch
from at_amd.workloads import da_polar_grid, error_realisation, large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SHAPES = [(64, 1, 1), (128, 1, 4), (256, 1, 2), (512, 1, 1), (64, 2, 1), (128, 2, 2),
          (256, 2, 2), (512, 2, 1)]
IDS = [f"b{b}p{p}u{u}" for b, p, u in SHAPES]


def _beam(n=1500):
    """A DA grid reaching past the apertures, with the special inputs."""
    r = da_polar_grid(n, xmax=12e-3, ymax=5e-3)
    s = zoo.beam(16, seed=3, scale=0.5)
    r[:, :8] = s[:, :8]
    return np.asfortranarray(r)


@pytest.fixture(scope="module")
def ring():
    return large_ring(ncells=2, cavity=True)


def _track(dl, ring, r0, nturns, shape, **kw):
    dl.set_launch_shape(*shape)
    r = r0.copy(order="F")
    out = dl.track(r, nturns, kw.pop("refpts", None), energy=ring.energy,
                   rest_energy=ring.particle.rest_energy, **kw)
    dl.set_launch_shape()
    return r, out


@pytest.mark.parametrize("shape", SHAPES, ids=IDS)
@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_shape_does_not_change_results(ring, shape, strict):
    dl = engine.DeviceLattice(ring)
    r0 = _beam()
    print(f"shape {shape} strict={strict}", flush=True)
    ref, (_, lm_ref) = _track(dl, ring, r0, 9, (0, 0, 0), losses=True, strict=strict)
    got, (_, lm) = _track(dl, ring, r0, 9, shape, losses=True, strict=strict)
    assert lm["islost"].sum() > 100
    np.testing.assert_array_equal(got, ref)
    for k in ("islost", "turn", "elem", "coord"):
        np.testing.assert_array_equal(lm[k], lm_ref[k], err_msg=k)
    # refpts: lost particles are not compacted away, their copies stay NaN
    refs = np.array([0, 7, len(ring)], np.uint32)
    a, ra = _track(dl, ring, r0, 3, (0, 0, 0), refpts=refs, strict=strict)
    b, rb = _track(dl, ring, r0, 3, shape, refpts=refs, strict=strict)
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(rb, ra)
    # no loss map: a NaN-x input keeps its other coordinates
    a, _ = _track(dl, ring, r0, 4, (0, 0, 0), strict=strict)
    b, _ = _track(dl, ring, r0, 4, shape, strict=strict)
    np.testing.assert_array_equal(b, a)
    assert np.isfinite(b[1, 0]) or not np.isfinite(r0[1, 0])
    if strict:
        o = r0.copy(order="F")
        _, lmo = po.atpass(ring, o, 9, losses=True)
        np.testing.assert_array_equal(got, o)
        for k in ("islost", "turn", "elem", "coord"):
            np.testing.assert_array_equal(lm[k], lmo[k], err_msg=k)


@pytest.mark.parametrize("shape", [(256, 2, 2), (64, 2, 1), (128, 1, 4)],
                         ids=["b256p2u2", "b64p2u1", "b128p1u4"])
def test_shape_in_repacked_batches(ring, shape):
    """Batched launches with survivors repacked between turn blocks (one
    graph), items of different turn counts: equal to the default plain
    launch."""
    lats = [error_realisation(ring, 700 + i) for i in range(3)]
    dls = [engine.DeviceLattice(l) for l in lats]
    parts = [da_polar_grid(n, xmax=12e-3, ymax=5e-3) for n in (900, 333, 1200)]
    turns = [(12, 0), (5, 2), (9, 1)]
    e, rest = ring.energy, ring.particle.rest_energy
    base = [device.DeviceArray.from_host(p.T) for p in parts]
    plain = [b.clone() for b in base]
    lm0 = track_batch(dls, plain, 0, energy=e, rest_energy=rest, item_turns=turns)
    shaped = [b.clone() for b in base]
    block, ppt, unroll = shape
    lm1 = track_batch(dls, shaped, 0, energy=e, rest_energy=rest, item_turns=turns,
                      block_size=block, ppt=ppt, unroll=unroll, repack=True)
    device.synchronize()
    nlost = 0
    for i in range(3):
        np.testing.assert_array_equal(shaped[i].to_host(), plain[i].to_host(), err_msg=str(i))
        a, b = lm1[i].to_host(), lm0[i].to_host()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{i} {k}")
        nlost += int(a["islost"].sum())
    assert nlost > 0

