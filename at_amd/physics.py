"""Closed orbit and transfer matrices by tracking on the engine.

Restatement of pyat's orbit and matrix finders (pyat/at/physics/orbit.py,
matrix.py, energy_loss.py): every one-turn map evaluation is a small batch
of particles (the reference orbit plus one displaced copy per coordinate)
tracked through the lattice in one engine launch.

* find_orbit4 / find_sync_orbit / find_orbit6 / find_orbit: Newton
  iterations on the one-turn map with forward-difference Jacobians
  (orbit.py:16-72 _orbit_dp, :74-120 _orbit_dct, :307-368 _orbit6);
* find_m44 / find_m66: central differences around the closed orbit
  (matrix.py:19-168); find_elem_m66: one element (matrix.py:171-205);
* get_energy_loss / get_timelag_fromU0 (energy_loss.py:31-206) for the
  synchronous phase that seeds find_orbit6.
"""

from __future__ import annotations

import math
import warnings
from enum import Enum

import numpy as np

from . import tracking
from .lattice import AtError, AtWarning, Dipole

CLIGHT = 299792458.0
# CODATA 2022: electron rest energy [eV] and classical electron radius [m];
# Cgamma = 4 pi r_e / (3 (m_e c^2)^3) [m/eV^3] (pyat/at/constants.py)
E_MASS = 510998.95069
E_RADIUS = 2.8179403205e-15
CGAMMA = 4.0 * math.pi * E_RADIUS / 3.0 / E_MASS ** 3


class DConstant:
    """Default numerical parameters (pyat/at/lattice/options.py:78-81)."""

    XYStep = 3.0e-8
    DPStep = 3.0e-6
    OrbConvergence = 1.0e-12
    OrbMaxIter = 20
    TStol = 1.0e-9


class ELossMethod(Enum):
    """energy_loss.py:19-29"""

    INTEGRAL = 1
    TRACKING = 2


# One-turn tracking in place through the engine (internal_lpass); returns
# the refpt copies (6, N, R, 1).
def _lpass(ring, r_in, refpts=None, keep_lattice=False):
    refs = np.zeros(0, np.uint32) if refpts is None else refpts
    return tracking.internal_lpass(ring, r_in, 1, refpts=refs, keep_lattice=keep_lattice)


def _epass(elem, r_in, **kwargs):
    return tracking.internal_epass(elem, r_in, **kwargs)


def _refs(ring, refpts) -> np.ndarray:
    return tracking.get_uint32_index(ring, refpts)


def _options(kwargs, *, dp_step=False):
    conv = kwargs.pop("convergence", DConstant.OrbConvergence)
    maxit = kwargs.pop("max_iterations", DConstant.OrbMaxIter)
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    dps = kwargs.pop("DPStep", DConstant.DPStep)
    return conv, maxit, xy, dps


def _newton(ring, ref_in, delta, rows, cols, ident, target, keep_lattice, conv, maxit):
    """Newton iteration on the one-turn map: the first columns of the
    tracked batch are ref_in + delta, the last one is ref_in itself."""
    nd = delta.shape[1]
    change = 1.0
    it = 0
    scaling = np.array([delta[i, i] for i in range(nd)])
    while change > conv and it < maxit:
        in_mat = np.asfortranarray(ref_in.reshape((6, 1)) + np.concatenate(
            (delta, np.zeros((6, 1))), axis=1))
        _lpass(ring, in_mat, keep_lattice=keep_lattice)
        ref_out = in_mat[:, nd]
        jac = (in_mat[np.ix_(rows, cols)] - in_mat[rows, nd:nd + 1]) / scaling[cols]
        a = jac - ident
        b = target(ref_out, ref_in)
        r_next = ref_in.copy()
        r_next[cols] = ref_in[cols] - np.linalg.solve(a, b)
        change = float(np.linalg.norm(r_next - ref_in))
        it += 1
        ref_in = r_next
        keep_lattice = True
    if it == maxit:
        warnings.warn(AtWarning("Maximum number of iterations reached. "
                                "Possible non-convergence"), stacklevel=3)
    return ref_in


def _orbit_dp(ring, dp=None, guess=None, keep_lattice=False, **kwargs):
    """Fixed momentum deviation (orbit.py:16-72)."""
    conv, maxit, xy, _ = _options(kwargs)
    if kwargs:
        raise AtError(f"Unexpected keywords for orbit_dp: {', '.join(kwargs)}")
    ref_in = np.zeros(6) if guess is None else np.array(guess, dtype=float)
    ref_in[4] = 0.0 if dp is None else dp
    delta = np.zeros((6, 4))
    for i in range(4):
        delta[i, i] = xy
    idx = np.arange(4)
    return _newton(ring, ref_in, delta, idx, idx, np.eye(4),
                   lambda out, rin: out[:4] - rin[:4], keep_lattice, conv, maxit)


def _orbit_dct(ring, dct=None, guess=None, keep_lattice=False, **kwargs):
    """Fixed path lengthening (orbit.py:74-120)."""
    conv, maxit, xy, _ = _options(kwargs)
    if kwargs:
        raise AtError(f"Unexpected keywords for orbit_dct: {', '.join(kwargs)}")
    ref_in = np.zeros(6) if guess is None else np.array(guess, dtype=float)
    delta = np.zeros((6, 5))
    for i in range(5):
        delta[i, i] = xy
    theta5 = np.zeros(5)
    theta5[4] = 0.0 if dct is None else dct
    id5 = np.zeros((5, 5))
    id5[:4, :4] = np.eye(4)
    rows = np.array([0, 1, 2, 3, 5])
    cols = np.arange(5)
    return _newton(ring, ref_in, delta, rows, cols, id5,
                   lambda out, rin: out[rows] - np.append(rin[:4], 0.0) - theta5,
                   keep_lattice, conv, maxit)


def _propagate(ring, orbit, refpts, keep_lattice):
    refs = _refs(ring, refpts)
    if refs.size == 0:
        return np.empty((0, 6))
    r = np.asfortranarray(orbit.reshape(6, 1).copy())
    out = _lpass(ring, r, refs, keep_lattice)
    return np.squeeze(out, axis=(1, 3)).T


def find_orbit4(ring, dp=None, refpts=None, *, dct=None, df=None, orbit=None,
                keep_lattice=False, **kwargs):
    """4D closed orbit for a given dp (orbit.py:123-212)."""
    if len([v for v in (dp, dct, df) if v is not None]) > 1:
        raise AtError("For off-momentum specification, only one of dp, dct and df may be "
                      "specified")
    if orbit is None:
        if df is not None:
            frf = cell_revolution_frequency(ring) * cell_harmnumber(ring)
            dct = -cell_length(ring) * df / (frf + df)
            orbit = _orbit_dct(ring, dct, keep_lattice=keep_lattice, **kwargs)
        elif dct is not None:
            orbit = _orbit_dct(ring, dct, keep_lattice=keep_lattice, **kwargs)
        else:
            orbit = _orbit_dp(ring, dp, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _propagate(ring, orbit, refpts, keep_lattice)


def find_sync_orbit(ring, dct=None, refpts=None, *, dp=None, df=None, orbit=None,
                    keep_lattice=False, **kwargs):
    """Closed orbit for a given path lengthening (orbit.py:215-304)."""
    if len([v for v in (dp, dct, df) if v is not None]) > 1:
        raise AtError("For off-momentum specification, only one of dp, dct and df may be "
                      "specified")
    if orbit is None:
        if df is not None:
            frf = cell_revolution_frequency(ring) * cell_harmnumber(ring)
            dct = -cell_length(ring) * df / (frf + df)
            orbit = _orbit_dct(ring, dct, keep_lattice=keep_lattice, **kwargs)
        elif dp is not None:
            orbit = _orbit_dp(ring, dp, keep_lattice=keep_lattice, **kwargs)
        else:
            orbit = _orbit_dct(ring, dct, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _propagate(ring, orbit, refpts, keep_lattice)


# --- ring quantities ---------------------------------------------------------

def cell_length(ring) -> float:
    return float(sum(getattr(e, "Length", 0.0) for e in ring))


def _gamma_beta(ring):
    rest = float(getattr(ring.particle, "rest_energy", 0.0))
    if rest == 0.0:
        return math.inf, 1.0
    gamma = ring.energy / rest
    return gamma, math.sqrt(1.0 - 1.0 / (gamma * gamma))


def beta(ring) -> float:
    return _gamma_beta(ring)[1]


def cell_revolution_frequency(ring) -> float:
    return beta(ring) * CLIGHT / cell_length(ring)


def _cavities(ring):
    return [e for e in ring if getattr(e, "PassMethod", "").endswith("CavityPass")]


def get_rf_frequency(ring) -> float:
    cav = _cavities(ring)
    if not cav:
        raise AtError("No cavity found in the lattice")
    freqs = np.unique([c.Frequency for c in cav])
    if freqs.size > 1:
        raise AtError("Cavities have different frequencies")
    return float(freqs[0])


def cell_harmnumber(ring) -> int:
    return int(round(get_rf_frequency(ring) / cell_revolution_frequency(ring)))


def is_6d(ring) -> bool:
    """Cavities active or radiating passes present (lattice_object.py)."""
    for e in ring:
        pm = getattr(e, "PassMethod", "")
        if pm.endswith("RadPass") or pm.endswith("CavityPass"):
            return True
    return False


def get_energy_loss(ring, method=ELossMethod.INTEGRAL) -> float:
    """Energy loss per turn [eV] (energy_loss.py:31-102); the integral
    counts dipoles only."""
    if method is ELossMethod.INTEGRAL:
        i2 = sum(e.BendingAngle ** 2 / e.Length for e in ring
                 if isinstance(e, Dipole) and e.Length > 0)
        coef = CGAMMA / 2.0 / math.pi * ring.energy ** 4
        return getattr(ring, "periodicity", 1) * coef * i2
    if method is ELossMethod.TRACKING:
        delta = 0.0
        for e in ring:
            if getattr(e, "PassMethod", "").endswith("RadPass"):
                r = np.zeros((6, 1), order="F")
                _epass(e, r, energy=ring.energy, particle=ring.particle)
                delta += r[4, 0]
        return -delta * ring.energy * getattr(ring, "periodicity", 1)
    raise AtError(f"Invalid method: {method}")


def get_timelag_fromU0(ring, *, method=ELossMethod.TRACKING, cavpts=None):
    """TimeLag that makes the synchronous phase zero, for cavities of one
    frequency and TimeLag (energy_loss.py:105-206, single-frequency branch)."""
    cav = _cavities(ring) if cavpts is None else [ring[i] for i in _refs(ring, cavpts)]
    u0 = get_energy_loss(ring, method=method) / getattr(ring, "periodicity", 1)
    freq = np.unique([c.Frequency for c in cav])
    tl0 = np.unique([getattr(c, "TimeLag", 0.0) for c in cav])
    if freq.size > 1 or tl0.size > 1:
        raise AtError("cavities of several frequencies or time lags are not supported")
    vrf = float(np.sum([c.Voltage for c in cav]))
    if u0 > vrf:
        per = getattr(ring, "periodicity", 1)
        raise AtError(f"The RF voltage ({per * vrf:.3e} eV) is lower than "
                      f"the radiation losses ({per * u0:.3e} eV).")
    timelag = CLIGHT / (2 * math.pi * float(freq[0])) * math.asin(u0 / vrf)
    return timelag * np.ones(len(cav)), timelag - float(tl0[0])


def _orbit6(ring, cavpts=None, guess=None, keep_lattice=False, **kwargs):
    """6D closed orbit (orbit.py:307-368)."""
    conv, maxit, xy, dps = _options(kwargs)
    method = kwargs.pop("method", ELossMethod.TRACKING)
    if kwargs:
        raise AtError(f"Unexpected keywords for orbit6: {', '.join(kwargs)}")
    l0 = cell_length(ring)
    f_rf = get_rf_frequency(ring)
    harm = round(f_rf * l0 / beta(ring) / CLIGHT)
    if guess is None:
        ref_in = np.zeros(6)
        _, dt = get_timelag_fromU0(ring, method=method, cavpts=cavpts)
        ref_in[5] = -dt
    else:
        ref_in = np.array(guess, dtype=float)
    theta = np.zeros(6)
    theta[5] = beta(ring) * CLIGHT * harm / f_rf - l0
    scaling = xy * np.array([1.0, 1.0, 1.0, 1.0, 0.0, 0.0]) + dps * np.array(
        [0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    delta = np.diag(scaling)
    idx = np.arange(6)
    return _newton(ring, ref_in, delta, idx, idx, np.eye(6),
                   lambda out, rin: out - rin - theta, keep_lattice, conv, maxit)


def find_orbit6(ring, refpts=None, *, orbit=None, keep_lattice=False, **kwargs):
    """6D closed orbit (orbit.py:371-458)."""
    if not _cavities(ring):
        raise AtError("No cavity found in the lattice")
    if orbit is None:
        orbit = _orbit6(ring, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _propagate(ring, orbit, refpts, keep_lattice)


def find_orbit(ring, refpts=None, **kwargs):
    """find_orbit6 for 6D lattices, find_sync_orbit when dct or df is given,
    find_orbit4 otherwise (orbit.py:461-519)."""
    if is_6d(ring):
        for k in ("dp", "dct", "df"):
            if kwargs.get(k) is not None:
                raise AtError(f"{k} is not supported on a 6D lattice")
            kwargs.pop(k, None)
        return find_orbit6(ring, refpts=refpts, **kwargs)
    if kwargs.get("dct") is not None or kwargs.get("df") is not None:
        return find_sync_orbit(ring, refpts=refpts, **kwargs)
    return find_orbit4(ring, refpts=refpts, **kwargs)


# --- transfer matrices --------------------------------------------------------

_JMT = np.array([[0.0, 1.0], [-1.0, 0.0]])
_JMT4 = np.kron(np.eye(2), _JMT)


def find_m44(ring, dp=None, refpts=None, dct=None, df=None, orbit=None,
             keep_lattice=False, **kwargs):
    """4x4 one-turn matrix by central differences (matrix.py:19-102)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    full = kwargs.pop("full", False)
    if orbit is None:
        orbit, _ = find_orbit4(ring, dp=dp, dct=dct, df=df, keep_lattice=keep_lattice, XYStep=xy)
        keep_lattice = True
    scaling = xy * np.ones(4)
    dg = np.concatenate((0.5 * np.diag(scaling), np.zeros((2, 4))))
    dmat = np.concatenate((dg, -dg), axis=1)
    in_mat = np.asfortranarray(orbit.reshape(6, 1) + dmat)
    refs = _refs(ring, refpts)
    out = _lpass(ring, in_mat, refs, keep_lattice)
    out_mat = np.rollaxis(np.squeeze(out, axis=3), -1)
    m44 = (in_mat[:4, :4] - in_mat[:4, 4:]) / scaling
    if refs.size > 0:
        mstack = (out_mat[:, :4, :4] - out_mat[:, :4, 4:]) / scaling
        if full:
            mstack = np.stack([m.dot(m44.dot(_JMT4.T.dot(m.T.dot(_JMT4)))) for m in mstack])
    else:
        mstack = np.empty((0, 4, 4))
    return m44, mstack


def find_m66(ring, refpts=None, orbit=None, keep_lattice=False, **kwargs):
    """6x6 one-turn matrix by central differences (matrix.py:105-168)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    dps = kwargs.pop("DPStep", DConstant.DPStep)
    if orbit is None:
        if is_6d(ring):
            orbit, _ = find_orbit6(ring, keep_lattice=keep_lattice, XYStep=xy, DPStep=dps,
                                   **kwargs)
        else:
            orbit, _ = find_orbit4(ring, keep_lattice=keep_lattice, XYStep=xy, **kwargs)
        keep_lattice = True
    scaling = xy * np.array([1.0, 1.0, 1.0, 1.0, 0.0, 0.0]) + dps * np.array(
        [0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    dg = 0.5 * np.diag(scaling)
    dmat = np.concatenate((dg, -dg), axis=1)
    in_mat = np.asfortranarray(orbit.reshape(6, 1) + dmat)
    refs = _refs(ring, refpts)
    out = _lpass(ring, in_mat, refs, keep_lattice)
    out_mat = np.rollaxis(np.squeeze(out, axis=3), -1)
    m66 = (in_mat[:, :6] - in_mat[:, 6:]) / scaling
    if refs.size > 0:
        mstack = (out_mat[:, :, :6] - out_mat[:, :, 6:]) / scaling
    else:
        mstack = np.empty((0, 6, 6))
    return m66, mstack


def find_elem_m66(elem, orbit=None, **kwargs):
    """6x6 matrix of one element (matrix.py:171-205)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    if orbit is None:
        orbit = np.zeros(6)
    scaling = xy * np.ones(6)
    dg = 0.5 * np.diag(scaling)
    dmat = np.concatenate((dg, -dg), axis=1)
    in_mat = np.asfortranarray(orbit.reshape(6, 1) + dmat)
    _epass(elem, in_mat, **kwargs)
    return (in_mat[:, :6] - in_mat[:, 6:]) / scaling
