"""GPU: the frozen orbit corrections of the large ring's error realisations
(at_amd/data/large_ring_corrections.npz, written on the engine by
tests/golden/make_corrections.py) reproduce their corrected closed orbits,
and the corrected rings keep particles that the uncorrected ones lose."""

import numpy as np
import pytest

from at_amd import workloads

pytestmark = pytest.mark.gpu


def test_frozen_settings_give_the_stored_orbits():
    c = workloads.frozen_corrections()
    ring = workloads.large_ring(radiation=False, cavity=False)
    bpms = workloads.correction_bpms(ring)
    n = len(bpms)
    for i in (0, 1, len(c["seeds"]) - 1):
        seed = int(c["seeds"][i])
        print(f"seed {seed}", flush=True)
        lat = workloads.error_realisation(ring, seed, float(c["shift"]), float(c["roll"]),
                                          float(c["gradient"]), settings=c["settings"][i])
        o, b = workloads._bpm_orbit(lat, bpms, None)
        assert o is not None, seed
        np.testing.assert_allclose([np.std(b[:n]), np.std(b[n:])], c["orbit_rms"][i], rtol=1e-6)
        # without the correction the full-size errors leave no closed orbit
        # or one an order of magnitude larger
        raw = workloads.error_realisation(ring, seed, float(c["shift"]), float(c["roll"]),
                                          float(c["gradient"]))
        o2, b2 = workloads._bpm_orbit(raw, bpms, None)
        assert o2 is None or np.std(b2[:n]) > 5 * np.std(b[:n])
