"""GPU: per-element diffusion matrices (atx_diffusion_matrices).

The device restates the reference's diff_* propagation B <- M B M^T + D of
the radiating cores (diff_drift.c, diff_str_kick.c, diff_bnd_kick.c,
diff_bend_fringe.c, diff_str_exactkick.c) and is built with the reference
operation order. Each element's matrix is compared with the reference
cores themselves, run from the same entrance orbit with a non-NULL bdiff
(at_diffmatrix, pyat/at.c:755-831): bit for bit, except where the
reference calls tan() of a momentum-dependent angle (bend fringe modes
1-3), where the device tan may differ by an ulp and the tolerance is
1e-12 of the matrix's largest entry."""

import numpy as np
import pytest

import zoo
from at_amd import engine
from at_amd.lattice import Lattice
from at_amd.workloads import large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

RAD = ("StrMPoleSymplectic4RadPass", "BndMPoleSymplectic4RadPass", "ExactMultipoleRadPass")
TAN = {"bend_rad_fringe_m1", "bend_rad_fringe_m2", "bend_rad_fringe_m3"}
CASES = [c for c in zoo.cases() if c[1].PassMethod in RAD]


def _compare(got, ref, exact, what):
    if exact:
        np.testing.assert_array_equal(got, ref, err_msg=what)
    else:
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale, err_msg=what)


@pytest.mark.parametrize("name,elem,energy", CASES, ids=[c[0] for c in CASES])
def test_element_matches_reference(name, elem, energy):
    ring = Lattice([elem], energy=energy)
    dl = engine.DeviceLattice(ring)
    rest = ring.particle.rest_energy
    rng = np.random.default_rng(11)
    for k in range(4):
        # the on-axis orbit gives zero diffusion in a pure quadrupole
        orbit = rng.normal(0.0, [1e-3, 1e-4, 5e-4, 5e-5, 2e-3, 1e-3]) if k else np.zeros(6)
        got = dl.diffusion_matrices(orbit.reshape(6, 1), energy, rest)[0]
        ref = po.diffusion_ref(elem, orbit, energy, ring.particle)
        if k:
            assert np.abs(ref).max() > 0.0
        _compare(got, ref, name not in TAN, f"{name} orbit {k}")


def test_non_radiating_elements_are_zero():
    elems = [c[1] for c in zoo.cases() if c[1].PassMethod not in RAD][:12]
    ring = Lattice(elems, energy=6e9)
    dl = engine.DeviceLattice(ring)
    orbits = np.tile(np.array([1e-4, 0, 2e-5, 0, 1e-4, 0.0])[:, None], (1, len(elems)))
    got = dl.diffusion_matrices(orbits, ring.energy, ring.particle.rest_energy)
    assert np.all(got == 0.0)


def test_ring_from_tracked_orbit():
    ring = large_ring(ncells=4, cavity=True)
    energy, rest = ring.energy, ring.particle.rest_energy
    dl = engine.DeviceLattice(ring)
    n = len(ring)
    r0 = np.asfortranarray(np.array([[2e-4], [1e-5], [1e-4], [-1e-5], [1e-3], [1e-4]]))
    rout = dl.track(r0.copy(order="F"), 1, np.arange(n, dtype=np.uint32), energy=energy,
                    rest_energy=rest, strict=True)
    orbits = np.asfortranarray(rout[:, 0, :, 0])
    got = dl.diffusion_matrices(orbits, energy, rest)
    nrad = 0
    for i, e in enumerate(ring):
        ref = po.diffusion_ref(e, orbits[:, i], energy, ring.particle)
        if e.PassMethod in RAD:
            nrad += 1
            _compare(got[i], ref, True, f"element {i} {e.PassMethod}")
        else:
            assert np.all(got[i] == 0.0) and np.all(ref == 0.0)
    assert nrad == 92  # 4 cells of 7 bends and 16 quadrupoles


def test_device_pointers_and_shape_errors():
    import torch

    ring = large_ring(ncells=1, cavity=False)
    dl = engine.DeviceLattice(ring)
    n = len(ring)
    with pytest.raises(ValueError):
        dl.diffusion_matrices(np.zeros((6, n + 1)), ring.energy, ring.particle.rest_energy)
    orbits = np.zeros((6, n), order="F")
    host = dl.diffusion_matrices(orbits, ring.energy, ring.particle.rest_energy)
    dev = torch.device("cuda:0")
    o = torch.zeros((n, 6), dtype=torch.float64, device=dev)
    b = torch.zeros((n, 36), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    engine._check(dl.lib.atx_diffusion_matrices(
        dl.h, o.data_ptr(), b.data_ptr(), float(ring.energy), float(ring.particle.rest_energy),
        stream.cuda_stream, engine.ATX_DEVICE_POINTERS), dl.lib)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.cpu().numpy().reshape(n, 6, 6).transpose(0, 2, 1), host)
