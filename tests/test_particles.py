"""CPU: particle distributions (at_amd.particles, pyat/at/tracking/
particles.py). The correlation matrices are pinned by their closed forms
(Twiss blocks times emittances), by the equilibrium envelope of the hmba
ring (whose mode emittances are the reference's MATLAB results, also used
in tests/test_envelope.py), and by the sample covariance of the drawn
particles; the ring cases track with the oracle (tests/oracle_hooks.py)."""

import os

import numpy as np
import pytest

from at_amd import envelope, load_mat
from at_amd.lattice import AtError
from at_amd.particles import beam, emittances_from_beam, sigma_matrix
from oracle import pyoracle as po

import oracle_hooks

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# pyat/test/test_physics.py:499-503 (MATLAB)
MODE_EMITTANCES = np.array([1.3203595267442241e-10, 1.1573828900796540e-37, 2.8586919200065274e-06])
from test_envelope import MODE_MATRICES  # noqa: E402  (the same MATLAB record)


def _uncoupled():
    return dict(alphax=0.3, betax=7.0, emitx=1.3e-10, alphay=-0.8, betay=2.5, emity=4e-12,
                blength=3e-3, espread=1e-3)


def test_uncoupled_closed_form():
    p = _uncoupled()
    s = sigma_matrix(**p)
    ex = p["emitx"] * np.array([[7.0, -0.3], [-0.3, (1 + 0.09) / 7.0]])
    ey = p["emity"] * np.array([[2.5, 0.8], [0.8, (1 + 0.64) / 2.5]])
    ez = 3e-6 * np.array([[1e-3 / 3e-3, 0.0], [0.0, 3e-3 / 1e-3]])
    np.testing.assert_allclose(s[:2, :2], ex, rtol=1e-15)
    np.testing.assert_allclose(s[2:4, 2:4], ey, rtol=1e-15)
    np.testing.assert_allclose(s[4:, 4:], ez, rtol=1e-15)
    assert np.all(s[:2, 2:] == 0.0) and np.all(s[:4, 4:] == 0.0)
    eig_em, proj = emittances_from_beam(sigma_mat=s)
    np.testing.assert_allclose(proj, [p["emitx"], p["emity"], 3e-6], rtol=1e-12)
    np.testing.assert_allclose(eig_em, proj, rtol=1e-9)


def test_missing_parameters_and_monochromatic_warning():
    with pytest.raises(AtError, match="missing"):
        sigma_matrix(alphax=0.0, betax=1.0)
    with pytest.raises(AtError, match="blength"):
        sigma_matrix(alphax=0.0, betax=1.0, alphay=0.0, betay=1.0, emitx=1.0, emity=1.0,
                     espread=1e-3)
    with pytest.warns(Warning, match="Monochromatic"):
        s = sigma_matrix(alphax=0.0, betax=1.0, alphay=0.0, betay=1.0, emitx=1.0, emity=1.0)
    assert np.all(s[4:, 4:] == 0.0)
    with pytest.raises(AtError):
        emittances_from_beam()


def test_beam_covariance_and_sigma_from_beam():
    s = sigma_matrix(**_uncoupled())
    # couple x and delta a little (dispersion-like correlation)
    s[0, 4] = s[4, 0] = 0.3 * np.sqrt(s[0, 0] * s[4, 4])
    orbit = np.array([1e-4, 0.0, -2e-5, 0.0, 1e-3, 0.0])
    b = beam(400_000, s, orbit=orbit, rng=np.random.default_rng(7))
    assert b.shape == (6, 400_000) and b.flags.f_contiguous
    np.testing.assert_allclose(b.mean(axis=1), orbit, atol=5e-3 * np.sqrt(np.diag(s)).max())
    got = sigma_matrix(beam=b)
    d = b - b.mean(axis=1)[:, None]
    np.testing.assert_allclose(got, d @ d.T / b.shape[1], rtol=1e-12, atol=0)
    scale = np.sqrt(np.outer(np.diag(s), np.diag(s)))
    assert np.abs((got - s) / scale).max() < 0.01
    em, _ = emittances_from_beam(beam=b)
    em_s, _ = emittances_from_beam(sigma_mat=s)
    np.testing.assert_allclose(np.sort(em), np.sort(em_s), rtol=0.02)


def test_beam_without_vertical_emittance():
    p = _uncoupled()
    p["emity"] = 0.0
    s = sigma_matrix(**p)
    b = beam(1000, s, rng=np.random.default_rng(1))
    assert np.all(b[2:4] == 0.0)
    assert np.all(np.isfinite(b)) and b[0].std() > 0 and b[4].std() > 0


@pytest.fixture
def hooks(monkeypatch):
    if not (po.available("oracle") and po.available("ref")):
        pytest.skip("oracle not built")
    oracle_hooks.install(monkeypatch)


def test_ring_6d_equilibrium(hooks):
    ring = load_mat(os.path.join(GOLDEN, "hmba.mat")).radiation_on(copy=True)
    s = sigma_matrix(ring)
    # the eigen-emittances are the ring's equilibrium ones (MATLAB) for x and
    # z; the vertical one is zero up to rounding
    eig_em, _ = emittances_from_beam(sigma_mat=s)
    assert abs(eig_em[0] / MODE_EMITTANCES[0] - 1) < 1e-6
    assert eig_em[1] < 1e-20
    assert abs(eig_em[2] / MODE_EMITTANCES[2] - 1) < 1e-6
    # Sigma = sum_k eps_k R_k with the MATLAB mode matrices and emittances
    ref = np.einsum("k,kij->ij", MODE_EMITTANCES, MODE_MATRICES)
    hz = [0, 1, 4, 5]
    sc = np.sqrt(np.outer(np.diag(ref), np.diag(ref)))[np.ix_(hz, hz)]
    assert np.abs((s - ref)[np.ix_(hz, hz)] / sc).max() < 1e-6
    # its beam sizes are the equilibrium envelope's (the envelope also holds
    # small correlations between the modes, which the sum leaves out)
    emit0, _, _ = envelope.ohmi_envelope(ring)
    d_env = np.diag(emit0["r66"])[hz]
    np.testing.assert_allclose(np.diag(s)[hz], d_env, rtol=1e-3)
    # given emittances replace the equilibrium ones
    s2 = sigma_matrix(ring, emitx=2 * MODE_EMITTANCES[0], emity=1e-12)
    eig2, _ = emittances_from_beam(sigma_mat=s2)
    np.testing.assert_allclose(eig2, [2 * MODE_EMITTANCES[0], 1e-12, MODE_EMITTANCES[2]],
                               rtol=1e-6)


def test_ring_4d_needs_emittances(hooks):
    ring = load_mat(os.path.join(GOLDEN, "hmba.mat")).radiation_off(copy=True)
    with pytest.raises(AtError, match="Emittances"):
        sigma_matrix(ring)
    with pytest.warns(Warning, match="Monochromatic"):
        s = sigma_matrix(ring, emitx=1e-10, emity=1e-12)
    # an uncoupled ring: each plane's block has determinant emit^2
    np.testing.assert_allclose(np.linalg.det(s[:2, :2]), 1e-20, rtol=1e-8)
    np.testing.assert_allclose(np.linalg.det(s[2:4, 2:4]), 1e-24, rtol=1e-8)
    # the horizontal block is emitx times the Twiss matrix of the periodic
    # solution: invariant under the one-turn map
    from at_amd import physics

    m44, _ = physics.find_m44(ring, 0.0)
    mx = m44[:2, :2]
    np.testing.assert_allclose(mx @ s[:2, :2] @ mx.T, s[:2, :2], rtol=1e-7, atol=1e-20)
