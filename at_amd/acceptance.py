"""Dynamic-aperture (acceptance) driver on the engine.

Restates pyat's boundary search (pyat/at/acceptance/boundary.py) and its
front-ends (acceptance.py): particle grids in one or two planes (RADIAL,
CARTESIAN), the survival mask of one multi-turn tracking launch
(get_survived: losses=True, refpts=None, in place), the boundary of the
surviving region, and the RECURSIVE radial search that tracks only the
particles next to the boundary. Every tracking call goes to the engine.
GridMode.FLOODFILL is not offered (the reference has it for CPU
multiprocessing only, acceptance.py:142-144).
"""

from __future__ import annotations

import warnings
from collections import namedtuple
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


_AXES = {"x": 0, "px": 1, "y": 2, "py": 3, "dp": 4, "ct": 5, "delta": 4, "s": 5,
         "xp": 1, "yp": 3}


def _axis_index(plane) -> int:
    if isinstance(plane, (int, np.integer)):
        return int(plane)
    key = str(plane)
    if key in _AXES:
        return _AXES[key]
    if key.lower() in _AXES:
        return _AXES[key.lower()]
    raise AtError(f"Unknown plane {plane!r}")


_Config = namedtuple("config", "planes planesi amplitudes shape bounds mode shift_zero")
_Grid = namedtuple("grid", "grid offset")


def grid_configuration(planes, npoints, amplitudes, grid_mode, bounds=None, shift_zero=1.0e-6):
    """boundary.py:70-109"""
    planes = np.atleast_1d(planes)
    ndims = len(planes)
    if ndims > 2 or ndims == 0:
        raise AtError("planes can have 1 or 2 element (1D or 2D aperture)")
    if ndims == 1 and grid_mode is GridMode.RADIAL:
        grid_mode = GridMode.CARTESIAN
    if np.shape(np.atleast_1d(npoints)) != (ndims,):
        raise AtError("npoints shape should be (len(planes),)")
    if np.shape(np.atleast_1d(amplitudes)) != (ndims,):
        raise AtError("amplitudes shape should be (len(planes),)")
    if bounds is not None and np.shape(np.atleast_2d(bounds)) != (ndims, 2):
        raise AtError("bounds shape should be (len(planes),2)")
    if grid_mode is GridMode.RADIAL or grid_mode is GridMode.RECURSIVE:
        if bounds is None:
            bounds = np.array([[0, 1], [np.pi, 0]])
        bounds = np.array(bounds, dtype=float)
        bounds[0][bounds[0] == 0] = 1.0e-6
    elif grid_mode is GridMode.CARTESIAN:
        if bounds is None:
            bounds = np.array([[p - 1, 1] for p in range(ndims)])
    else:
        raise AtError(f"GridMode {grid_mode} undefined.")
    bounds = tuple(map(tuple, np.atleast_2d(bounds)))
    return _Config(planes, np.array([_axis_index(p) for p in planes]),
                   np.atleast_1d(amplitudes), np.atleast_1d(np.array(npoints, dtype=np.int32)),
                   bounds, grid_mode, shift_zero)


def get_parts(config, offset):
    """(6, n) particles of the grid around ``offset`` (boundary.py:112-148)."""
    amp, npt, bnd = config.amplitudes, config.shape, config.bounds
    if config.mode is GridMode.CARTESIAN:
        x = [np.linspace(*b, n) * a for a, b, n in zip(amp, bnd, npt)]
        try:
            g1, g2 = np.meshgrid(*x)
        except ValueError:
            g1, g2 = np.meshgrid(x, 0.0)
        g = np.array([g1.flatten(), g2.flatten()])
    elif config.mode is GridMode.RADIAL:
        x = [np.linspace(*b, n) for b, n in zip(bnd, npt)]
        g1, g2 = np.meshgrid(*x)
        g = np.array([(amp[0] * np.cos(g2) * g1).flatten(), (amp[1] * np.sin(g2) * g1).flatten()])
    else:
        raise AtError(f"GridMode {config.mode} has no fixed grid")
    parts = np.zeros((6, int(np.prod(npt))))
    parts[config.planesi, :] = [g[i] for i in range(len(config.planesi))]
    offset = np.array(offset) + config.shift_zero
    parts = (parts.T + offset).T
    return np.asfortranarray(parts), _Grid(g, offset[config.planesi])


# Tracking hooks (the engine): survival after nturns, and projection of
# particles from an observation point to the end of the ring.
def _survived(parts, ring, nturns, **kwargs):
    """boundary.py:151-163"""
    _, _, td = tracking.lattice_track(ring, parts, nturns=nturns, losses=True, refpts=None,
                                      in_place=True, **kwargs)
    return np.invert(td["loss_map"].islost)


def _project(ring, parts, obs, **kwargs):
    tracking.lattice_track(ring, parts, 1, refpts=None, in_place=True, start_elem=int(obs),
                           **kwargs)


def set_ring_orbit(ring, dp, obspt, orbit):
    """Initial closed orbit and rotated ring (boundary.py:57-67)."""
    if obspt is not None:
        if not np.isscalar(obspt):
            raise AtError("Scalar value needed for obspt")
        ringrot = ring.rotate(int(obspt))
    else:
        ringrot = ring
    if orbit is None:
        orbit0, orbitobs = physics.find_orbit(ring, dp=dp, refpts=obspt)
        orbit = orbit0 if obspt is None else np.squeeze(orbitobs)
    return np.asarray(orbit, dtype=float), ringrot


def _nearest_order(grid):
    """Order boundary points by angle, then by nearest neighbour on the
    normalised grid (boundary.py:170-214)."""
    angle = np.arctan2(*grid)
    norm = np.linalg.norm(grid.T, axis=1)
    val, inv = np.unique(angle, return_inverse=True)
    gf = np.zeros((2, len(val)))
    for i in range(len(val)):
        inds = np.where(inv == i)[0]
        ni = norm[inds]
        gf[:, i] = grid[:, inds[np.where(ni == np.amax(ni))[0]][0]]
    gf = gf[:, np.argsort(np.arctan2(*gf))]
    x, y = gf[0, :].copy(), gf[1, :].copy()
    dxmin = min(np.diff(np.unique(x)))
    dymin = min(np.diff(np.unique(y)))
    iorder = [0]
    for i in range(1, len(x)):
        xnow, ynow = x[iorder[-1]], y[iorder[-1]]
        dd = np.sqrt(((x - xnow) / dxmin) ** 2 + ((y - ynow) / dymin) ** 2)
        if i <= 3:
            ic = [j for j in np.argsort(dd) if j not in iorder]
        else:
            direction = np.sign(iorder[-1] - iorder[-2])
            ic = [j for j in np.argsort(dd)
                  if j not in iorder and np.sign(j - iorder[-1]) == direction]
        if ic:
            iorder.append(ic[0])
    xnow, ynow = x[iorder[-1]], y[iorder[-1]]
    dd = np.sqrt(((x[iorder[0]] - xnow) / dxmin) ** 2 + ((y[iorder[0]] - ynow) / dymin) ** 2)
    if dd < 1.5:
        iorder.append(iorder[0])
    return gf[:, iorder]


def _search_bnd(ma, sa):
    bnd = np.zeros((2, 1))
    for j, m in enumerate(ma):
        bnd = sa[:, j]
        if not m and j > 0:
            bnd = sa[:, j - 1]
            break
    return bnd


def get_grid_boundary(mask, grid, config):
    """Boundary of the surviving part of a grid (boundary.py:166-272)."""
    from scipy.ndimage import binary_dilation, binary_opening

    if not np.any(mask):
        warnings.warn(physics.AtWarning("No particle survived, please check your grid or "
                                        "lattice. Acceptance set to [0.0, 0.0]."),
                      stacklevel=2)
        return np.zeros(2) if len(config.shape) == 1 else np.zeros((2, 1))
    if config.mode is GridMode.RADIAL:
        angles = np.round(np.arctan2(*grid.grid), decimals=8)
        angles, invi = np.unique(angles, return_inverse=True)
        bnd = np.zeros((2, len(angles)))
        for i in range(len(angles)):
            bnd[:, i] = _search_bnd(mask[invi == i], grid.grid[:, invi == i])
        return bnd
    if config.mode is GridMode.CARTESIAN:
        cnt = np.flip(config.shape)
        if len(cnt) == 1:
            g = grid.grid
            xp, xn = g[0] >= 0, g[0] <= 0
            bp = _search_bnd(mask[xp], g[:, xp])
            bn = _search_bnd(np.flip(mask[xn]), np.flip(g[:, xn], axis=1))
            return np.squeeze([bn[0], bp[0]])
        mask2d = np.reshape(mask.copy(), cnt)
        bnd1 = binary_opening(mask2d, np.ones((2, 1)))
        bnd2 = binary_opening(mask2d, np.ones((1, 2)))
        bnd = np.logical_and(bnd1, bnd2)
        bnd = np.logical_and(binary_dilation(bnd == 0, border_value=1), bnd)
        return _nearest_order(grid.grid[:, bnd.reshape(mask.shape)])
    raise AtError(f"GridMode {config.mode} undefined.")


def grid_boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None, dp=None,
                         offset=None, bounds=None, grid_mode=GridMode.RADIAL, verbose=False,
                         shift_zero=1.0e-9, **kwargs):
    """Track whole grids, one launch for all observation points
    (boundary.py:275-404, FLOODFILL excluded)."""
    config = grid_configuration(planes, npoints, amplitudes, grid_mode, bounds=bounds,
                                shift_zero=shift_zero)
    if offset is None:
        offset = [None for _ in range(len(obspt))]
    allparts, grids = [], []
    for obs, orbit in zip(obspt, offset):
        orbit, _ = set_ring_orbit(ring, dp, obs, orbit)
        parts, grid = get_parts(config, orbit)
        # project particles from the observation point to the ring's end
        _project(ring, parts, 0 if obs is None else obs, **kwargs)
        allparts.append(parts)
        grids.append(grid)
    allparts = np.asfortranarray(np.concatenate(allparts, axis=1))
    mask = np.split(_survived(allparts, ring, nturns, **kwargs), len(grids))
    survived = [g.grid[:, m] for g, m in zip(grids, mask)]
    boundary = [get_grid_boundary(m, g, config) for g, m in zip(grids, mask)]
    grids = [g.grid for g in grids]
    if len(obspt) == 1:
        return boundary[0], survived[0], grids[0]
    return boundary, survived, grids


def recursive_boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None,
                              dp=None, offset=None, bounds=None, divider=2, verbose=False,
                              shift_zero=1.0e-9, **kwargs):
    """Radial search along each angle, halving the step after a loss
    (boundary.py:407-503). Each round tracks the new trial points of all
    angles in one launch."""
    offset, newring = set_ring_orbit(ring, dp, obspt, offset)
    config = grid_configuration(planes, npoints, amplitudes, GridMode.RECURSIVE, bounds=bounds,
                                shift_zero=shift_zero)
    rtol = min(np.atleast_1d(config.amplitudes / config.shape))
    rsteps = config.amplitudes
    if len(np.atleast_1d(config.shape)) == 2:
        angles = np.linspace(*config.bounds[1], config.shape[1])
    else:
        angles = np.linspace(*config.bounds[1], 2)
    angles = np.atleast_1d(angles)
    planesi = config.planesi

    ftol = min(rtol / rsteps)
    cs = np.squeeze([np.cos(angles), np.sin(angles)])
    cs = np.around(cs, decimals=9)
    fact = np.ones(len(angles))
    survived = np.full(len(angles), True)
    part = np.zeros((6, len(angles)))
    grid = np.array([])
    mask = np.array([])
    while np.any(survived):
        for i, pi in enumerate(planesi):
            part[pi, survived] += cs[i, survived] * rsteps[i] * fact[survived]
        istracked = np.array([not np.any([np.allclose(p, g, rtol=1.0e-9) for g in grid.T])
                              for p in part[planesi].T])
        survived = np.array([np.any([np.allclose(p, m, rtol=1.0e-9) for m in mask.T])
                             for p in part[planesi].T])
        pt = part[:, istracked]
        grid = np.hstack([grid, pt[planesi]]) if grid.size else pt[planesi]
        ptmp = np.asfortranarray((pt.T + offset).T)
        if ptmp.shape[1] > 0:
            survived[istracked] = _survived(ptmp, newring, nturns, **kwargs)
        pm = part[:, np.logical_and(istracked, survived)]
        mask = np.hstack([mask, pm[planesi]]) if mask.size else pm[planesi]
        for i in range(len(angles)):
            if not survived[i] and fact[i] > ftol:
                deltas = cs[:, i] * rsteps[:] * min(1, 2 * fact[i])
                if np.any(abs(deltas) > abs(part[planesi, i])):
                    part[planesi, i] = np.zeros(len(planesi))
                else:
                    for j, pi in enumerate(planesi):
                        part[pi, i] -= deltas[j]
                survived[i] = True
                fact[i] *= 1 / divider
    for i, pi in enumerate(planesi):
        part[pi] -= cs[i] * rsteps[i] * fact
    return np.squeeze(part[planesi]), mask, grid


def boundary_search(ring, planes, npoints, amplitudes, nturns=1024, obspt=None, dp=None,
                    offset=None, bounds=None, grid_mode=GridMode.RADIAL, use_mp=False,
                    verbose=False, shift_zero=1.0e-9, **kwargs):
    """Loss boundary at one or more observation points (boundary.py:506-600)."""
    rp = (tracking.get_uint32_index(ring, obspt) if obspt is not None
          else np.atleast_1d(obspt))
    if offset is not None:
        try:
            offset = np.broadcast_to(offset, (len(rp), 6))
        except ValueError:
            raise AtError(f"offset and refpts have incoherent shapes: {np.shape(offset)}, "
                          f"{np.shape(obspt)}") from None
    divider = kwargs.pop("divider", 2)
    kwargs.pop("gpu_pool", None)
    kwargs.pop("pool_size", None)
    kwargs.pop("start_method", None)
    if grid_mode is GridMode.FLOODFILL:
        raise AtError("floodfill is not implemented for GPU tracking")
    if grid_mode is GridMode.RECURSIVE:
        if offset is None:
            offset = [None for _ in rp]
        out = [recursive_boundary_search(ring, planes, npoints, amplitudes, nturns=nturns,
                                         obspt=r, dp=dp, offset=o, bounds=bounds,
                                         divider=divider, verbose=verbose,
                                         shift_zero=shift_zero, **kwargs)
               for r, o in zip(rp, offset)]
        if len(rp) == 1:
            return out[0]
        return tuple(list(x) for x in zip(*out))
    return grid_boundary_search(ring, planes, npoints, amplitudes, nturns=nturns, obspt=rp,
                                dp=dp, offset=offset, bounds=bounds, grid_mode=grid_mode,
                                verbose=verbose, shift_zero=shift_zero, **kwargs)


def get_acceptance(ring, planes, npoints, amplitudes, nturns=1024, refpts=None, dp=None,
                   offset=None, bounds=None, grid_mode=GridMode.RADIAL, use_mp=False,
                   gpu_pool=None, verbose=False, divider=2, shift_zero=1.0e-6,
                   start_method=None, **kwargs):
    """Acceptance at the ``refpts`` observation points (acceptance.py:27-199).
    Returns (boundary, survived, tracked)."""
    axes = list(np.atleast_1d(planes))
    for i, a in enumerate(axes):
        if a in ("xp", "yp"):
            warnings.warn(physics.AtWarning(f"Axis name {a} is deprecated."), stacklevel=2)
            axes[i] = a[::-1]
    planes = tuple(axes) if len(axes) > 1 else axes[0]
    if gpu_pool is not None:
        tracking.engine.set_device(int(gpu_pool[0]))
    return boundary_search(ring, planes, npoints, amplitudes, nturns=nturns, obspt=refpts,
                           dp=dp, offset=offset, bounds=bounds, grid_mode=grid_mode,
                           use_mp=use_mp, verbose=verbose, divider=divider,
                           shift_zero=shift_zero, **kwargs)


def get_1d_acceptance(ring, plane, resolution, amplitude, nturns=1024, refpts=None, dp=None,
                      offset=None, grid_mode=GridMode.RADIAL, use_mp=False, gpu_pool=None,
                      verbose=False, divider=2, shift_zero=1.0e-6, start_method=None, **kwargs):
    """1D acceptance (acceptance.py:202-303): without multiprocessing the
    reference forces the RECURSIVE search, and so does this driver."""
    if not use_mp:
        grid_mode = GridMode.RECURSIVE
    if len(np.atleast_1d(plane)) != 1:
        raise AtError("1D acceptance: single plane required")
    if not (np.isscalar(resolution) and np.isscalar(amplitude)):
        raise AtError("1D acceptance: scalar args required")
    npoint = np.ceil(amplitude / resolution)
    if grid_mode is not GridMode.RECURSIVE and not npoint > 1:
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
