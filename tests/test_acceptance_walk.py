"""CPU: the RECURSIVE radial walk of at_amd.acceptance on synthetic
survival functions (no tracking), for the properties that fix its answer:

* monotone survival (alive inside radius R): the boundary is the last
  multiple of the finest step below R, on every ray;
* the walk's look-ahead changes only the number of launches: the boundary,
  the tracked list and the survived list (with their order) are identical
  with and without it, and no point is tracked twice by the walk;
* with a hole in the survival region the answer is a surviving point whose
  next finest step outwards is lost.
"""

import numpy as np
import pytest

from at_amd import acceptance
from at_amd.acceptance import GridMode, _layout, _Walk


class FakeTracker:
    """Survival of particles by their (x, y) position: inside an ellipse of
    semi-axes ``rx``, ``ry`` and outside an optional annular hole."""

    def __init__(self, rx, ry, hole=None):
        self.rx, self.ry, self.hole = rx, ry, hole
        self.launches = 0
        self.seen = []

    def __call__(self, parts, ring, nturns, **kwargs):
        self.launches += 1
        x, y = parts[0], parts[2]
        rho = np.hypot(x / self.rx, y / self.ry)
        alive = rho < 1.0
        if self.hole is not None:
            lo, hi = self.hole
            alive &= ~((rho > lo) & (rho < hi))
        self.seen.append(parts[[0, 2]].copy())
        return alive


def _walk(tracker, monkeypatch, lookahead, npoints=(10, 7), amps=(10e-3, 6e-3), planes=("x", "y")):
    monkeypatch.setattr(acceptance, "_survived", tracker)
    lay = _layout(list(planes), list(npoints), list(amps), GridMode.RECURSIVE)
    return _Walk(lay, 2, np.zeros(6)).run(None, 16, lookahead, {})


@pytest.mark.parametrize("lookahead", [0, 32])
def test_monotone_boundary_is_last_finest_multiple(monkeypatch, lookahead):
    rx, ry = 7.3e-3, 4.1e-3
    edge, kept, tracked = _walk(FakeTracker(rx, ry), monkeypatch, lookahead)
    # finest step: amplitude / 16 (npoints 10 -> tolerance 1/10 -> 4 halvings)
    unit = 1.0 / 16
    lay = _layout(["x", "y"], [10, 7], [10e-3, 6e-3], GridMode.RECURSIVE)
    angles = np.linspace(*lay.bounds[1], 7)
    cs = np.around([np.cos(angles), np.sin(angles)], 9)
    for k in range(angles.size):
        d = np.array([cs[0, k] * 10e-3, cs[1, k] * 6e-3])
        frac = 1.0 / np.hypot(d[0] / rx, d[1] / ry)          # boundary in units of the amplitude
        expect = (np.ceil(frac / unit - 1e-12) - 1) * unit
        np.testing.assert_allclose(edge[:, k], d * expect, rtol=0, atol=1e-15)
    assert kept.shape[0] == 2 and tracked.shape[1] >= kept.shape[1]


@pytest.mark.parametrize("hole", [None, (0.35, 0.45)])
def test_lookahead_changes_launches_only(monkeypatch, hole):
    a, b = FakeTracker(7.3e-3, 4.1e-3, hole), FakeTracker(7.3e-3, 4.1e-3, hole)
    e0, k0, t0 = _walk(a, monkeypatch, 0)
    e4, k4, t4 = _walk(b, monkeypatch, 32)
    np.testing.assert_array_equal(e0, e4)
    np.testing.assert_array_equal(k0, k4)
    np.testing.assert_array_equal(t0, t4)
    assert b.launches < a.launches
    # without look-ahead every launch holds only points the walk visits, once each
    visited = np.concatenate(a.seen, axis=1)
    assert visited.shape[1] == t0.shape[1]
    assert np.unique(visited.round(15), axis=1).shape[1] == visited.shape[1]


def test_hole_answer_is_a_survivor_next_to_a_loss(monkeypatch):
    tr = FakeTracker(7.3e-3, 4.1e-3, (0.35, 0.45))
    edge, kept, tracked = _walk(tr, monkeypatch, 32)
    alive = tr(np.vstack([edge[0], 0 * edge[0], edge[1], 0 * edge[0], 0 * edge[0],
                          0 * edge[0]]), None, 0)
    lay = _layout(["x", "y"], [10, 7], [10e-3, 6e-3], GridMode.RECURSIVE)
    angles = np.linspace(*lay.bounds[1], 7)
    cs = np.around([np.cos(angles), np.sin(angles)], 9)
    step = np.array([cs[0] * 10e-3, cs[1] * 6e-3]) / 16
    beyond = edge + step
    lost = ~tr(np.vstack([beyond[0], 0 * beyond[0], beyond[1], 0 * beyond[0], 0 * beyond[0],
                          0 * beyond[0]]), None, 0)
    at_zero = np.all(np.abs(edge) < 1e-15, axis=0)
    assert np.all((alive | at_zero) & lost)


def test_one_plane_walk_both_directions(monkeypatch):
    tr = FakeTracker(2.45e-3, 1.0)
    monkeypatch.setattr(acceptance, "_survived", tr)
    lay = _layout(["x"], [1], [1e-3], GridMode.RECURSIVE)
    edge, kept, tracked = _Walk(lay, 2, np.zeros(6)).run(None, 16, 32, {})
    # npoints 1: the step is the whole amplitude and is never halved
    np.testing.assert_allclose(edge, [-2e-3, 2e-3], rtol=0, atol=1e-15)
    assert tracked.shape == (1, 6)
