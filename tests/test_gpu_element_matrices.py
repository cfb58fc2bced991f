"""GPU: atx_element_matrices, the 6x6 matrix of every element of a ring
around its entrance orbit in one launch, against find_elem_m66 element by
element on the engine (pyat/at/physics/matrix.py:171-202, called per
element by _dmatr, pyat/at/physics/radiation.py:78-90) and against the
oracle's element passes of the same probes."""

import os

import numpy as np
import pytest

from at_amd import engine, load_mat, physics
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEP = physics.DConstant.XYStep


@pytest.fixture(scope="module")
def ring():
    """hmba with radiation on and its elements' entrance orbits: the 6-D
    closed orbit propagated to every element."""
    r = load_mat(os.path.join(GOLDEN, "hmba.mat")).radiation_on(copy=True)
    _, orbits = physics.find_orbit6(r, refpts=range(len(r)))
    return r, np.asfortranarray(orbits.T)


def oracle_matrix(elem, orbit, energy, particle):
    """matrix.py:195-202 with the oracle's element pass."""
    d = 0.5 * STEP * np.eye(6)
    probes = np.asfortranarray(orbit.reshape(6, 1) + np.concatenate((d, -d), axis=1))
    po.elempass(elem, probes, energy=energy, particle=particle)
    return (probes[:, :6] - probes[:, 6:]) / STEP


@pytest.mark.parametrize("strict", (True, False))
def test_one_launch_equals_element_by_element(ring, strict):
    r, orbits = ring
    print(f"hmba: {len(r)} elements, strict={strict}", flush=True)
    ms = physics.element_matrices(r, orbits, strict=strict)
    assert ms.shape == (len(r), 6, 6)
    for i in range(len(r)):
        one = engine.DeviceLattice([r[i]])
        try:
            probes = np.asfortranarray(orbits[:, i:i + 1]
                                       + np.concatenate((0.5 * STEP * np.eye(6),
                                                         -0.5 * STEP * np.eye(6)), axis=1))
            one.elempass(0, probes, energy=r.energy, rest_energy=r.particle.rest_energy,
                         strict=strict)
        finally:
            one.close()
        np.testing.assert_array_equal(ms[i], (probes[:, :6] - probes[:, 6:]) / STEP,
                                      err_msg=f"element {i} {r[i].FamName}")


@pytest.mark.skipif(not po.available("oracle"), reason="oracle not built")
def test_strict_matrices_equal_the_oracle(ring):
    r, orbits = ring
    ms = physics.element_matrices(r, orbits, strict=True)
    for i in range(len(r)):
        want = oracle_matrix(r[i], orbits[:, i], r.energy, r.particle)
        np.testing.assert_array_equal(ms[i], want, err_msg=f"element {i} {r[i].FamName}")


def test_default_orbit_and_known_answer(ring):
    """Zero entrance orbit by default; element 19 of hmba as in
    pyat/test/test_physics.py:117-133; drifts are [[1, L], [0, 1]] blocks."""
    r, _ = ring
    ms = physics.element_matrices(r)
    elem19 = np.array([[1.0386, 0.180911, 0.0, 0.0, 0.0, 0.0],
                       [0.434959, 1.0386, 0.0, 0.0, 0.0, 0.0],
                       [0.0, 0.0, 0.961891, 0.176344, 0.0, 0.0],
                       [0.0, 0.0, -0.423978, 0.961891, 0.0, 0.0],
                       [0.0, 0.0, 0.0, 0.0, 1.0, 0.0],
                       [0.0, 0.0, 0.0, 0.0, 0.0, 1.0]])
    np.testing.assert_allclose(ms[19], elem19, rtol=1e-5, atol=1e-7)
    ndrift = 0
    for i, e in enumerate(r):
        if e.PassMethod == "DriftPass":
            ndrift += 1
            np.testing.assert_allclose(ms[i][:4, :4][0, 1], e.Length, rtol=1e-9)
            np.testing.assert_allclose(np.diag(ms[i]), np.ones(6), rtol=0, atol=1e-12)
    assert ndrift > 10


def test_device_pointers_and_argument_errors(ring):
    """ATX_DEVICE_POINTERS gives the host path's matrices; a zero step and
    NULL arrays are value errors."""
    from at_amd import device

    r, orbits = ring
    dl = engine.DeviceLattice(list(r))
    try:
        host = dl.element_matrices(orbits, STEP, r.energy, r.particle.rest_energy)
        d_o = device.DeviceArray.from_host(np.ascontiguousarray(orbits.T))
        d_m = device.DeviceArray.zeros(36 * len(r))
        lib = engine.load_library()
        s = device.Stream()
        engine._check(lib.atx_element_matrices(dl.h, d_o.ptr, STEP, d_m.ptr, r.energy,
                                               r.particle.rest_energy, s.handle,
                                               engine.ATX_DEVICE_POINTERS), lib)
        s.synchronize()
        dev = d_m.to_host().reshape(len(r), 6, 6).transpose(0, 2, 1)
        np.testing.assert_array_equal(dev, host)
        assert lib.atx_element_matrices(dl.h, d_o.ptr, 0.0, d_m.ptr, r.energy, 0.0, None,
                                        engine.ATX_DEVICE_POINTERS) == engine.ATX_E_VALUE
        assert lib.atx_element_matrices(dl.h, None, STEP, None, r.energy, 0.0, None,
                                        0) == engine.ATX_E_VALUE
        d_o.close()
        d_m.close()
    finally:
        dl.close()
