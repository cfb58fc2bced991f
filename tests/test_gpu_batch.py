"""GPU: batched error realisations in one launch.

Each instance of a batch must give exactly what the same instance gives
alone (same kernel arithmetic, bit-exact in both builds) and, in the strict
build, what the oracle gives (bit-exact, RF cavity included)."""

import numpy as np
import pytest

from at_amd import device, engine
from at_amd.device import DeviceArray
from at_amd.autotune import tune_batch
from at_amd.batch import LossMap, track_batch
from at_amd.workloads import da_polar_grid, error_realisation, large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _instances(n_inst=3, ncells=2, cavity=True):
    base = large_ring(ncells=ncells, cavity=cavity)
    sizes = [300, 517, 64][:n_inst]
    lats = [error_realisation(base, 100 + i) for i in range(n_inst)]
    parts = [da_polar_grid(n, xmax=8e-3, ymax=4e-3) for n in sizes]
    return lats, parts


@pytest.mark.parametrize("strict", [True, False], ids=["strict", "fast"])
def test_batch_equals_single_and_oracle(strict):

    lats, parts = _instances()
    dls = [engine.DeviceLattice(l) for l in lats]
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    tens = [DeviceArray.from_host(p.T) for p in parts]
    lms = track_batch(dls, tens, 7, turn=2, energy=energy, rest_energy=rest, strict=strict,
                      block_size=128)
    device.synchronize()
    for i, (lat, p) in enumerate(zip(lats, parts)):
        got = tens[i].to_host().T
        single = p.copy(order="F")
        _, lm1 = dls[i].track(single, 7, None, turn=2, energy=energy, rest_energy=rest,
                              losses=True, strict=strict)
        np.testing.assert_array_equal(got, single)
        np.testing.assert_array_equal(lms[i].flag.to_host().astype(bool), lm1["islost"])
        np.testing.assert_array_equal(lms[i].turn.to_host(), lm1["turn"])
        np.testing.assert_array_equal(lms[i].elem.to_host(), lm1["elem"])
        ref = p.copy(order="F")
        _, lmo = po.atpass(lat, ref, 7, turn=2, losses=True)
        np.testing.assert_array_equal(lm1["islost"], lmo["islost"])
        np.testing.assert_array_equal(lm1["elem"], lmo["elem"])
        if strict:
            np.testing.assert_array_equal(single, ref)
        else:
            surv = ~lmo["islost"]
            np.testing.assert_allclose(single[:, surv], ref[:, surv], rtol=0, atol=1e-11)


def test_autotuner_picks_a_candidate_and_keeps_state(tmp_path, monkeypatch):

    monkeypatch.setenv("ATX_TUNE_CACHE", str(tmp_path / "tune.json"))
    lats, parts = _instances(cavity=True)
    dls = [engine.DeviceLattice(l) for l in lats]
    tens = [DeviceArray.from_host(p.T) for p in parts]
    before = [t.clone() for t in tens]
    res = tune_batch(dls, tens, energy=lats[0].energy, rest_energy=lats[0].particle.rest_energy,
                     use_cache=False, nturns=40)
    assert res["block_size"] in (64, 128, 256, 512)
    assert (res["ppt"], res["unroll"]) in {(1, 1), (1, 2), (1, 4), (2, 1), (2, 2)}
    # staged search: the five kernels at 256 lanes, then the other sizes
    assert len(res["ms"]) == 8 and f"b{res['block_size']}p{res['ppt']}u{res['unroll']}" in res["ms"]
    assert res["ms"][f"b{res['block_size']}p{res['ppt']}u{res['unroll']}"] == min(res["ms"].values())
    for a, b in zip(tens, before):
        np.testing.assert_array_equal(a.to_host(), b.to_host())


def _lm_arrays(lm):
    return (lm.flag.to_host().astype(bool), lm.turn.to_host(),
            lm.elem.to_host(), lm.coord.to_host())


@pytest.mark.parametrize("losses", [True, False], ids=["losses", "nolosses"])
@pytest.mark.parametrize("block", [64, 256])
def test_repacked_launch_equals_plain(losses, block):
    """Turn blocks with the survivors repacked between them (one graph):
    coordinates and loss maps equal the plain launch bit for bit, on every
    call (the graph is rebuilt and replayed per call), with per-item turn
    counts that end in different blocks and an item (third of five) that
    loses every particle on its first turn, so that later blocks map no
    workgroup to it or to the finished items."""

    base = large_ring(ncells=2, cavity=True)
    lats = [error_realisation(base, 300 + i) for i in range(5)]
    dls = [engine.DeviceLattice(l) for l in lats]
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    sizes = [2000, 700, 600, 1500, 333]
    parts = [da_polar_grid(n, xmax=12e-3, ymax=5e-3) for n in sizes]
    gone = np.zeros((6, 600))
    gone[0] = np.linspace(0.05, 0.08, 600)  # outside the 10 mm apertures
    parts[2] = np.asfortranarray(gone)
    turns = [(37, 0), (64, 5), (40, 0), (9, 1), (1, 0)]
    plain = [DeviceArray.from_host(p.T) for p in parts]
    lm_plain = track_batch(dls, plain, 0, energy=energy, rest_energy=rest, losses=losses,
                           block_size=block, item_turns=turns)
    device.synchronize()
    nlost = 0
    for call in range(3):
        rep = [DeviceArray.from_host(p.T) for p in parts]
        lm_rep = track_batch(dls, rep, 0, energy=energy, rest_energy=rest, losses=losses,
                             block_size=block, item_turns=turns, repack=True)
        # eager work between the graph runs
        device.synchronize()
        for i in range(len(parts)):
            a, b = rep[i].to_host(), plain[i].to_host()
            np.testing.assert_array_equal(a, b, err_msg=f"call {call} item {i}")
            if losses:
                for x, y in zip(_lm_arrays(lm_rep[i]), _lm_arrays(lm_plain[i])):
                    np.testing.assert_array_equal(x, y, err_msg=f"call {call} item {i}")
                nlost += int(lm_rep[i].flag.to_host().sum())
            else:
                nlost += int(np.isnan(a[:, 0]).sum())
    assert nlost > 0


def test_repacked_launch_continues_losses():
    """Two repacked calls with ATX_CONTINUE_LOSSES equal one plain call."""

    lats = [error_realisation(large_ring(ncells=2, cavity=True), 410 + i) for i in range(2)]
    dls = [engine.DeviceLattice(l) for l in lats]
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    parts = [da_polar_grid(900, xmax=12e-3, ymax=5e-3), da_polar_grid(450, xmax=12e-3, ymax=5e-3)]
    one = [DeviceArray.from_host(p.T) for p in parts]
    lm_one = track_batch(dls, one, 20, energy=energy, rest_energy=rest)
    two = [DeviceArray.from_host(p.T) for p in parts]
    lm_two = [LossMap.empty(p.shape[1]) for p in parts]
    track_batch(dls, two, 6, energy=energy, rest_energy=rest, loss_maps=lm_two, repack=True)
    track_batch(dls, two, 14, turn=6, energy=energy, rest_energy=rest, loss_maps=lm_two,
                continue_losses=True, repack=True)
    device.synchronize()
    for i in range(2):
        np.testing.assert_array_equal(two[i].to_host(), one[i].to_host())
        for x, y in zip(_lm_arrays(lm_two[i]), _lm_arrays(lm_one[i])):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("repack", [False, True], ids=["plain", "repacked"])
def test_equal_turn_batch_in_xcd_order_equals_single_launches(repack):
    """A batch whose instances all run the same turns is launched with its
    workgroups dealt to the XCDs in contiguous chunks (a workgroup-index
    map in the kernel, over the launch's workgroups or the live ones of
    each repacked turn block). Eleven instances of sizes from part of one
    workgroup to several, so that the chunks split instances and the
    workgroup counts are not multiples of the XCD count: every instance
    equals its own single-lattice launch bit for bit (coordinates and loss
    map)."""

    base = large_ring(ncells=2, cavity=True)
    sizes = [40, 900, 64, 333, 1500, 65, 700, 128, 2100, 5, 450]
    lats = [error_realisation(base, 500 + i) for i in range(len(sizes))]
    dls = [engine.DeviceLattice(l) for l in lats]
    energy, rest = lats[0].energy, lats[0].particle.rest_energy
    parts = [da_polar_grid(n, xmax=12e-3, ymax=5e-3) for n in sizes]
    nturns = 24
    print(f"xcd-order batch: {len(sizes)} instances, repack={repack}", flush=True)
    tens = [DeviceArray.from_host(p.T) for p in parts]
    lms = track_batch(dls, tens, nturns, energy=energy, rest_energy=rest, block_size=64,
                      repack=repack)
    device.synchronize()
    nlost = 0
    for i, p in enumerate(parts):
        single = p.copy(order="F")
        _, lm1 = dls[i].track(single, nturns, None, energy=energy, rest_energy=rest, losses=True)
        np.testing.assert_array_equal(tens[i].to_host().T, single, err_msg=f"item {i}")
        flag, turn, elem, coord = _lm_arrays(lms[i])
        np.testing.assert_array_equal(flag, lm1["islost"], err_msg=f"item {i}")
        np.testing.assert_array_equal(turn, lm1["turn"], err_msg=f"item {i}")
        np.testing.assert_array_equal(elem, lm1["elem"], err_msg=f"item {i}")
        nlost += int(flag.sum())
    assert nlost > 0
