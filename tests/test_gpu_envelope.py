"""GPU: the equilibrium envelope and the ring's particle distribution on the engine (diffusion matrices from
atx_diffusion_matrices, orbits and element matrices from the HIP tracking)
against the MATLAB known answers of tests/test_envelope.py, and the
cumulative diffusion against the same computation with the reference's
radiating cores and the oracle."""

import numpy as np
import pytest

from at_amd import envelope

import oracle_hooks
from test_envelope import hmba6, test_ohmi_envelope_known_answers  # noqa: F401
from test_particles import test_ring_4d_needs_emittances, test_ring_6d_equilibrium  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture
def hooks():
    """No hooks: every tracking and diffusion call goes to the engine."""
    return None


def test_cumulative_diffusion_matches_reference(hmba6, monkeypatch):  # noqa: F811
    orbit = np.array([-2.6352679435692906e-09, -1.4584877344311391e-10, 0.0, 0.0,
                      -6.6969442207863660e-06, -5.8845211247215173e-02])
    b_gpu, o_gpu = envelope.diffusion_matrices_along(hmba6, orbit=orbit)
    with monkeypatch.context() as m:
        oracle_hooks.install(m)
        b_ref, o_ref = envelope.diffusion_matrices_along(hmba6, orbit=orbit)
    np.testing.assert_allclose(o_gpu, o_ref, rtol=0, atol=1e-15)
    scale = np.abs(b_ref[-1]).max()
    np.testing.assert_allclose(b_gpu, b_ref, rtol=0, atol=1e-9 * scale)
