"""Dynamic-aperture (acceptance) driver on the engine.

The caller-facing surface is pyat's (pyat/at/acceptance/acceptance.py:
``get_acceptance``, ``get_1d_acceptance`` and its three plane shortcuts,
``GridMode``), returning (boundary, survived, tracked). Two ways of finding
the boundary:

* **Whole grids** (RADIAL, CARTESIAN; boundary.py:112-148 define the point
  sets): every point of every observation point's grid is moved to the
  ring's entrance (one partial turn from the observation point) and all of
  them are tracked in ONE launch; the survival mask is then read as an
  image (angle x radius, or y x x) and its edge extracted with array
  operations.
* **Radial walk** (RECURSIVE, boundary.py:409-503): along each angle a walk
  steps outwards; after a loss it steps back and continues at a finer step,
  until a loss at the finest step. The walk's points and results are the
  reference's (the boundary is quantised by the walk, and a point near the
  boundary may be chaotic, so only the same points give the same answer,
  bit for bit): the walk keeps each ray's position both as an integer
  number of finest steps (the identity of a point, for the cache of points
  already tracked) and as the floating-point coordinates the reference's
  arithmetic gives along the same path (the particles actually tracked).
  What is the engine's own is the scheduling: a round in which some ray
  needs a new point tracks, in the same launch, up to ``lookahead`` points
  that ray may need next — the walk played forward from the pending point
  for both outcomes of every undecided point, nearest decisions first —
  with their coordinates as each path would produce them. A speculative
  result is used only when the walk reaches the point along that path, so
  the tracked and survived lists, and the boundary, are the reference's,
  with several times fewer launches (a launch of a few hundred particles
  costs as much as one of a single particle).

GridMode.FLOODFILL is not offered (the reference runs it on CPU
multiprocessing only, acceptance.py:142-144).
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass
from enum import Enum

import numpy as np

from . import physics, tracking
from .lattice import AtError

__all__ = ["GridMode", "boundary_search", "get_acceptance", "get_1d_acceptance",
           "get_horizontal_acceptance", "get_vertical_acceptance", "get_momentum_acceptance"]


class GridMode(Enum):
    """boundary.py:23-31"""

    RADIAL = 0
    CARTESIAN = 1
    RECURSIVE = 2
    FLOODFILL = 3


_PLANE_INDEX = {"x": 0, "px": 1, "xp": 1, "y": 2, "py": 3, "yp": 3, "dp": 4, "delta": 4,
                "ct": 5, "s": 5}


def _plane_index(plane) -> int:
    if isinstance(plane, (int, np.integer)):
        return int(plane)
    idx = _PLANE_INDEX.get(str(plane), _PLANE_INDEX.get(str(plane).lower()))
    if idx is None:
        raise AtError(f"Unknown plane {plane!r}")
    return idx


@dataclass(frozen=True)
class _Layout:
    """What is searched: planes (names and coordinate indices), amplitude
    and point count per plane, bounds, mode and the shift off zero."""

    planes: np.ndarray
    index: np.ndarray
    amplitudes: np.ndarray
    counts: np.ndarray
    bounds: tuple
    mode: GridMode
    shift_zero: float


def _layout(planes, npoints, amplitudes, mode, bounds=None, shift_zero=1.0e-6) -> _Layout:
    """Validated search layout; RADIAL and RECURSIVE bounds are
    (radius range, angle range) with a zero radius moved to 1e-6,
    CARTESIAN bounds are per-plane multiples of the amplitude
    (boundary.py:70-109)."""
    planes = np.atleast_1d(planes)
    nd = planes.size
    if nd not in (1, 2):
        raise AtError("planes can have 1 or 2 element (1D or 2D aperture)")
    if nd == 1 and mode is GridMode.RADIAL:
        mode = GridMode.CARTESIAN
    counts = np.atleast_1d(np.array(npoints, dtype=np.int32))
    amps = np.atleast_1d(np.array(amplitudes, dtype=float))
    if counts.shape != (nd,):
        raise AtError("npoints shape should be (len(planes),)")
    if amps.shape != (nd,):
        raise AtError("amplitudes shape should be (len(planes),)")
    if bounds is not None and np.shape(np.atleast_2d(bounds)) != (nd, 2):
        raise AtError("bounds shape should be (len(planes),2)")
    if mode in (GridMode.RADIAL, GridMode.RECURSIVE):
        b = np.array([[0.0, 1.0], [np.pi, 0.0]] if bounds is None else bounds, dtype=float)
        b[0][b[0] == 0] = 1.0e-6
    elif mode is GridMode.CARTESIAN:
        b = np.array([[-1.0, 1.0], [0.0, 1.0]][:nd] if bounds is None else bounds, dtype=float)
    else:
        raise AtError(f"GridMode {mode} undefined.")
    return _Layout(planes, np.array([_plane_index(p) for p in planes]), amps, counts,
                   tuple(tuple(r) for r in np.atleast_2d(b)), mode, shift_zero)


# --- tracking hooks (the engine) ------------------------------------------------

def _survived(parts, ring, nturns, **kwargs):
    """Survival of each particle after ``nturns`` (one call, in place; a
    grid loses most of its points in the first turns, so the survivors are
    repacked between turn blocks)."""
    kwargs.setdefault("repack", True)
    _, _, td = tracking.lattice_track(ring, parts, nturns=nturns, losses=True, refpts=None,
                                      in_place=True, **kwargs)
    return np.invert(td["loss_map"].islost)


def _project(ring, parts, obs, **kwargs):
    """Carry particles from element ``obs`` to the end of the ring."""
    tracking.lattice_track(ring, parts, 1, refpts=None, in_place=True, start_elem=int(obs),
                           **kwargs)


def _orbit_and_ring(ring, dp, obspt, orbit):
    """Closed orbit at the observation point (unless given) and the ring
    starting there (boundary.py:57-67)."""
    start = ring
    if obspt is not None:
        if not np.isscalar(obspt):
            raise AtError("Scalar value needed for obspt")
        start = ring.rotate(int(obspt))
    if orbit is None:
        at_start, at_obs = physics.find_orbit(ring, dp=dp, refpts=obspt)
        orbit = at_start if obspt is None else at_obs[0]
    return np.asarray(orbit, dtype=float), start


# --- whole grids -------------------------------------------------------------------

def _grid_points(lay: _Layout) -> np.ndarray:
    """Grid coordinates, (2, n), in the reference's order and arithmetic:
    RADIAL radius-fastest over angles, CARTESIAN first-plane-fastest; a
    1-D grid carries a zero second row."""
    if lay.mode is GridMode.RADIAL:
        radius = np.linspace(*lay.bounds[0], lay.counts[0])
        angle = np.linspace(*lay.bounds[1], lay.counts[1])
        r = np.tile(radius, angle.size)
        a = np.repeat(angle, radius.size)
        return np.array([lay.amplitudes[0] * np.cos(a) * r, lay.amplitudes[1] * np.sin(a) * r])
    if lay.mode is GridMode.CARTESIAN:
        axes = [np.linspace(*b, n) * amp
                for amp, b, n in zip(lay.amplitudes, lay.bounds, lay.counts)]
        if len(axes) == 1:
            return np.array([axes[0], np.zeros_like(axes[0])])
        return np.array([np.tile(axes[0], axes[1].size), np.repeat(axes[1], axes[0].size)])
    raise AtError(f"GridMode {lay.mode} has no fixed grid")


def _grid_beam(lay: _Layout, points: np.ndarray, orbit: np.ndarray) -> np.ndarray:
    """(6, n) particles: the grid in its planes, plus the orbit shifted off
    zero by ``shift_zero`` on every coordinate."""
    beam = np.zeros((6, points.shape[1]))
    beam[lay.index] = points[:lay.index.size]
    return np.asfortranarray(beam + (orbit + lay.shift_zero)[:, None])


def _last_before_first_loss(alive: np.ndarray) -> np.ndarray:
    """For each row of a survival image ordered outwards, the column of
    the last point before the first loss, the innermost point not counted
    as a loss (it is the starting point of the row); the last column when
    nothing further out is lost."""
    n = alive.shape[1]
    lost = ~alive[:, 1:]
    first = np.where(lost.any(axis=1), lost.argmax(axis=1) + 1, n)
    return first - 1


def _cartesian_edge(alive: np.ndarray, points: np.ndarray) -> np.ndarray:
    """Edge of the surviving region of a 2-D cartesian image (rows = second
    plane): survivors with a surviving neighbour both along x and along y
    (strands one point wide are dropped) that touch a non-member or the
    grid's border; one point per direction (the farthest), by angle."""
    left = np.zeros_like(alive)
    left[:, 1:] = alive[:, :-1]
    right = np.zeros_like(alive)
    right[:, :-1] = alive[:, 1:]
    down = np.zeros_like(alive)
    down[1:, :] = alive[:-1, :]
    up = np.zeros_like(alive)
    up[:-1, :] = alive[1:, :]
    member = alive & (left | right) & (down | up)
    padded = np.pad(member, 1, constant_values=False)
    inner = member.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            inner &= padded[1 + dy:1 + dy + member.shape[0], 1 + dx:1 + dx + member.shape[1]]
    edge = (member & ~inner).reshape(-1)
    pts = points[:, edge]
    if pts.shape[1] == 0:
        return pts
    angle = np.round(np.arctan2(pts[1], pts[0]), 12)
    radius = np.hypot(pts[0], pts[1])
    order = np.lexsort((-radius, angle))
    angle, pts = angle[order], pts[:, order]
    keep = np.ones(angle.size, bool)
    keep[1:] = angle[1:] != angle[:-1]
    return pts[:, keep]


def _grid_boundary(alive: np.ndarray, points: np.ndarray, lay: _Layout):
    if not alive.any():
        warnings.warn(physics.AtWarning("No particle survived, please check your grid or "
                                        "lattice. Acceptance set to [0.0, 0.0]."), stacklevel=3)
        return np.zeros(2) if lay.counts.size == 1 else np.zeros((2, 1))
    if lay.mode is GridMode.RADIAL:
        # one ray per direction, keyed (and ordered) as the reference keys
        # them: arctan2(x, y) rounded to 8 decimals; rays that share a key
        # are searched as one, in point order
        nr, na = lay.counts
        image = alive.reshape(na, nr)
        rays = points.reshape(2, na, nr)
        key = np.round(np.arctan2(rays[0, :, 0], rays[1, :, 0]), 8)
        edge = []
        for k in np.unique(key):
            rows = np.flatnonzero(key == k)
            a = image[rows].reshape(1, -1)
            pts = rays[:, rows].reshape(2, -1)
            edge.append(pts[:, _last_before_first_loss(a)[0]])
        return np.array(edge).T
    if lay.counts.size == 1:
        x = points[0]
        outward = np.flatnonzero(x >= 0)
        inward = np.flatnonzero(x <= 0)[::-1]
        hi = outward[_last_before_first_loss(alive[outward][None, :])[0]]
        lo = inward[_last_before_first_loss(alive[inward][None, :])[0]]
        return np.array([x[lo], x[hi]])
    nx, ny = lay.counts
    return _cartesian_edge(alive.reshape(ny, nx), points)


def grid_boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None, dp=None,
                         offset=None, bounds=None, grid_mode=GridMode.RADIAL, verbose=False,
                         shift_zero=1.0e-9, **kwargs):
    """Boundary from whole grids at one or more observation points, all
    tracked in one launch (boundary.py:275-404 without FLOODFILL)."""
    lay = _layout(planes, npoints, amplitudes, grid_mode, bounds, shift_zero)
    points = _grid_points(lay)
    orbits = [None] * len(obspt) if offset is None else offset
    beams = []
    for obs, orbit in zip(obspt, orbits):
        orbit, _ = _orbit_and_ring(ring, dp, obs, orbit)
        beam = _grid_beam(lay, points, orbit)
        _project(ring, beam, 0 if obs is None else obs, **kwargs)
        beams.append(beam)
    alive = _survived(np.asfortranarray(np.concatenate(beams, axis=1)), ring, nturns, **kwargs)
    alive = alive.reshape(len(beams), points.shape[1])
    boundary = [_grid_boundary(a, points, lay) for a in alive]
    survived = [points[:, a] for a in alive]
    if len(beams) == 1:
        return boundary[0], survived[0], points
    return boundary, survived, [points] * len(beams)


# --- radial walk -------------------------------------------------------------------

class _Walk:
    """The RECURSIVE search along every ray at once.

    Per ray: ``pos`` (integer, in finest steps), ``level`` (0 = coarsest;
    the step at level l is divider**(depth - l) finest steps, the first
    being the whole amplitude), ``xyz`` (the floating-point coordinates in
    the search planes, advanced with the reference's operations:
    ``x += base*fact`` out, ``x -= base*min(1, 2*fact)`` back, or reset to
    zero when stepping back would cross it), ``done``."""

    def __init__(self, lay: _Layout, divider, offset):
        nd = lay.index.size
        rtol = np.min(lay.amplitudes / lay.counts)
        ftol = np.min(rtol / lay.amplitudes)
        facts = [1.0]
        while facts[-1] > ftol:
            facts.append(facts[-1] * (1 / divider))
        self.facts = facts                     # fact per level, as the reference computes it
        self.depth = len(facts) - 1            # the finest level
        self.unit = [divider ** (self.depth - lvl) for lvl in range(self.depth + 1)]
        angles = np.atleast_1d(np.linspace(*lay.bounds[1], lay.counts[1] if nd == 2 else 2))
        cs = np.around(np.array([np.cos(angles), np.sin(angles)]), decimals=9)
        # per ray: the coarse step in the search planes, and its components
        # along both directions used by the step-back test
        self.base = cs[:nd] * lay.amplitudes[:, None]
        self.test = cs * lay.amplitudes[:, None] if nd == 2 else cs * lay.amplitudes[0]
        self.nd, self.nray = nd, angles.size
        self.lay, self.offset = lay, offset
        self.pos = np.zeros(self.nray, np.int64)
        self.level = np.zeros(self.nray, np.int64)
        self.xyz = np.zeros((nd, self.nray))
        self.done = np.zeros(self.nray, bool)
        self.outcome = {}     # (ray, pos) -> survived, fixed at the first visit
        self.spare = {}       # (ray, pos, coordinate bytes) -> survived, tracked ahead

    # the reference's floating-point moves
    def _out(self, ray, xyz, level):
        return xyz + self.base[:, ray] * self.facts[level]

    def _back(self, ray, xyz, pos, level):
        m = min(1, 2 * self.facts[level])
        if np.any(np.abs(self.test[:, ray] * m) > np.abs(xyz)):
            return np.zeros_like(xyz), 0
        return xyz - self.base[:, ray] * m, pos - min(self.unit[0], 2 * self.unit[level])

    def _ahead(self, ray, xyz, pos, level, budget):
        """Up to ``budget`` points the ray may need after the pending one
        (xyz, pos, level), nearest decisions first: the walk is played
        forward from it for both outcomes of every point whose outcome is
        not known (assuming, along each branch, the outcomes that branch
        took), with the coordinates each path produces."""
        found, listed = [], {(pos, xyz.tobytes())}
        queue = [(xyz, pos, level, {})]
        expansions = 0
        while queue and len(found) < budget and expansions < 8 * budget:
            x, p, lvl, assumed = queue.pop(0)
            expansions += 1
            known = self.outcome.get((ray, p), assumed.get(p))
            for alive in ((known,) if known is not None else (True, False)):
                branch = dict(assumed)
                branch[p] = alive
                if alive:
                    nxt = (self._out(ray, x, lvl), p + self.unit[lvl], lvl)
                elif lvl == self.depth:
                    continue
                else:
                    xb, pb = self._back(ray, x, p, lvl)
                    nxt = (self._out(ray, xb, lvl + 1), pb + self.unit[lvl + 1], lvl + 1)
                nx, np_, nl = nxt
                tag = (np_, nx.tobytes())
                if ((ray, np_) not in self.outcome and np_ not in branch
                        and tag not in listed):
                    listed.add(tag)
                    found.append((nx, np_))
                queue.append((nx, np_, nl, branch))
        return found[:budget]

    def _track(self, wanted, ring, nturns, kwargs):
        beam = np.zeros((6, len(wanted)))
        for j, (_, xyz, _) in enumerate(wanted):
            beam[self.lay.index, j] = xyz
        beam = np.asfortranarray(beam + self.offset[:, None])
        alive = _survived(beam, ring, nturns, **kwargs)
        for (ray, xyz, p), a in zip(wanted, alive):
            self.spare[(ray, p, xyz.tobytes())] = bool(a)

    def run(self, ring, nturns, lookahead, kwargs):
        tracked, kept = [], []
        while not self.done.all():
            active = np.flatnonzero(~self.done)
            for r in active:
                self.xyz[:, r] = self._out(r, self.xyz[:, r], self.level[r])
                self.pos[r] += self.unit[self.level[r]]
            # rays whose point is new and not tracked ahead along this path
            need = [r for r in active if (r, self.pos[r]) not in self.outcome
                    and (r, self.pos[r], self.xyz[:, r].tobytes()) not in self.spare]
            if need:
                wanted = []
                for r in need:
                    wanted.append((r, self.xyz[:, r].copy(), self.pos[r]))
                    wanted += [(r, x, p) for x, p in
                               self._ahead(r, self.xyz[:, r], self.pos[r], self.level[r],
                                           lookahead)]
                self._track(wanted, ring, nturns, kwargs)
            for r in active:
                key = (r, self.pos[r])
                if key not in self.outcome:
                    alive = self.spare[(r, self.pos[r], self.xyz[:, r].tobytes())]
                    self.outcome[key] = alive
                    tracked.append(self.xyz[:, r].copy())
                    if alive:
                        kept.append(self.xyz[:, r].copy())
                if self.outcome[key]:
                    continue
                if self.level[r] == self.depth:
                    self.done[r] = True
                else:
                    self.xyz[:, r], self.pos[r] = self._back(r, self.xyz[:, r], self.pos[r],
                                                              self.level[r])
                    self.level[r] += 1
        # the boundary: one finest step inside the final loss
        edge = np.stack([self.xyz[:, r] - self.base[:, r] * self.facts[self.depth]
                         for r in range(self.nray)], axis=1)
        as_array = (lambda pts: np.array(pts).T if pts else np.zeros((self.nd, 0)))
        return np.squeeze(edge), as_array(kept), as_array(tracked)


def recursive_boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None,
                              dp=None, offset=None, bounds=None, divider=2, verbose=False,
                              shift_zero=1.0e-9, lookahead=32, **kwargs):
    """RECURSIVE boundary at one observation point (boundary.py:407-503)."""
    orbit, start = _orbit_and_ring(ring, dp, obspt, offset)
    lay = _layout(planes, npoints, amplitudes, GridMode.RECURSIVE, bounds, shift_zero)
    return _Walk(lay, divider, orbit).run(start, nturns, lookahead, kwargs)


def boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None, dp=None,
                    offset=None, bounds=None, grid_mode=GridMode.RADIAL, use_mp=False,
                    verbose=False, shift_zero=1.0e-9, **kwargs):
    """Loss boundary at one or more observation points (boundary.py:506-600)."""
    obs = (tracking.get_uint32_index(ring, obspt) if obspt is not None
           else np.atleast_1d(obspt))
    if offset is not None:
        try:
            offset = np.broadcast_to(offset, (len(obs), 6))
        except ValueError:
            raise AtError(f"offset and refpts have incoherent shapes: {np.shape(offset)}, "
                          f"{np.shape(obspt)}") from None
    divider = kwargs.pop("divider", 2)
    for k in ("gpu_pool", "pool_size", "start_method"):
        kwargs.pop(k, None)
    if grid_mode is GridMode.FLOODFILL:
        raise AtError("floodfill is not implemented for GPU tracking")
    if grid_mode is not GridMode.RECURSIVE:
        return grid_boundary_search(ring, planes, npoints, amplitudes, nturns=nturns,
                                    obspt=obs, dp=dp, offset=offset, bounds=bounds,
                                    grid_mode=grid_mode, verbose=verbose,
                                    shift_zero=shift_zero, **kwargs)
    orbits = [None] * len(obs) if offset is None else offset
    results = [recursive_boundary_search(ring, planes, npoints, amplitudes, nturns=nturns,
                                         obspt=o, dp=dp, offset=orb, bounds=bounds,
                                         divider=divider, shift_zero=shift_zero, **kwargs)
               for o, orb in zip(obs, orbits)]
    if len(results) == 1:
        return results[0]
    return tuple(list(parts) for parts in zip(*results))


def get_acceptance(ring, planes, npoints, amplitudes, nturns=1024, refpts=None, dp=None,
                   offset=None, bounds=None, grid_mode=GridMode.RADIAL, use_mp=False,
                   gpu_pool=None, verbose=False, divider=2, shift_zero=1.0e-6,
                   start_method=None, **kwargs):
    """Acceptance at the ``refpts`` observation points (acceptance.py:27-199).
    Returns (boundary, survived, tracked)."""
    names = []
    for p in np.atleast_1d(planes):
        if p in ("xp", "yp"):
            warnings.warn(physics.AtWarning(f"Axis name {p} is deprecated."), stacklevel=2)
            p = p[::-1]
        names.append(p)
    if gpu_pool is not None:
        tracking.engine.set_device(int(gpu_pool[0]))
    return boundary_search(ring, tuple(names) if len(names) > 1 else names[0], npoints,
                           amplitudes, nturns=nturns, obspt=refpts, dp=dp, offset=offset,
                           bounds=bounds, grid_mode=grid_mode, use_mp=use_mp, verbose=verbose,
                           divider=divider, shift_zero=shift_zero, **kwargs)


def get_1d_acceptance(ring, plane, resolution, amplitude, nturns=1024, refpts=None, dp=None,
                      offset=None, grid_mode=GridMode.RADIAL, use_mp=False, gpu_pool=None,
                      verbose=False, divider=2, shift_zero=1.0e-6, start_method=None, **kwargs):
    """Acceptance along one plane (acceptance.py:202-303). Without
    ``use_mp`` the reference runs the RECURSIVE walk whatever grid_mode
    says, and so does this driver."""
    if not use_mp:
        grid_mode = GridMode.RECURSIVE
    if np.size(plane) != 1:
        raise AtError("1D acceptance: single plane required")
    if not (np.isscalar(resolution) and np.isscalar(amplitude)):
        raise AtError("1D acceptance: scalar args required")
    npoint = np.ceil(amplitude / resolution)
    if grid_mode is not GridMode.RECURSIVE and npoint <= 1:
        raise AtError("Grid has only one point: increase amplitude or reduce resolution")
    b, s, g = get_acceptance(ring, plane, npoint, amplitude, nturns=nturns, dp=dp,
                             refpts=refpts, grid_mode=grid_mode, use_mp=use_mp,
                             gpu_pool=gpu_pool, verbose=verbose, divider=divider,
                             shift_zero=shift_zero, offset=offset, **kwargs)
    return np.squeeze(b), s, g


def get_horizontal_acceptance(ring, resolution, amplitude, *args, **kwargs):
    return get_1d_acceptance(ring, "x", resolution, amplitude, *args, **kwargs)


def get_vertical_acceptance(ring, resolution, amplitude, *args, **kwargs):
    return get_1d_acceptance(ring, "y", resolution, amplitude, *args, **kwargs)


def get_momentum_acceptance(ring, resolution, amplitude, *args, **kwargs):
    return get_1d_acceptance(ring, "dp", resolution, amplitude, *args, **kwargs)
