"""Particle distributions for tracking (pyat/at/tracking/particles.py):
the 6x6 correlation matrix of a beam (``sigma_matrix``), Gaussian
particles drawn from it (``beam``) and the emittances of a distribution
(``emittances_from_beam``). The generated (6, N) arrays are Fortran-ordered,
the layout the tracking entry points take.

For a ring the correlation matrix is assembled from the normal-mode
R-matrices of the one-turn map at the ring's start, weighted by the mode
emittances: for a 6D ring those of the equilibrium envelope
(``envelope.ohmi_envelope``), for a 4D ring the caller's emitx/emity with
the R-matrices of ``find_m44``.
"""

from __future__ import annotations

from warnings import warn

import numpy as np
from numpy.linalg import LinAlgError, cholesky, det
from scipy.linalg import block_diag, eig

from .lattice import AtError, AtWarning

__all__ = ["beam", "sigma_matrix", "emittances_from_beam"]

_J2 = np.array([[0.0, 1.0], [-1.0, 0.0]])
_JMAT = block_diag(_J2, _J2, _J2)


def _twiss_block(alpha: float, beta: float) -> np.ndarray:
    return np.array([[beta, -alpha], [-alpha, (1.0 + alpha * alpha) / beta]])


def sigma_matrix(ring=None, **kwargs) -> np.ndarray:
    """(6, 6) correlation matrix for particle generation
    (particles.py:20-178).

    Keywords: ``beam`` (a (6, N) distribution: all else is ignored),
    ``twiss_in`` (a record with ``R``, or ``alpha`` and ``beta``; needs
    emitx and emity), ``alphax``, ``betax``, ``emitx``, ``alphay``,
    ``betay``, ``emity``, ``blength`` and ``espread`` (both or neither: a
    monochromatic beam otherwise), ``orbit``. With a 6D ring the equilibrium
    emittances are used unless emitx/emity are given; a 4D ring needs them.
    """
    kw = dict(kwargs)

    def require(mess, *keys):
        miss = [k for k in keys if k not in kw]
        if miss:
            raise AtError(f"{mess}: missing {', '.join(miss)}")
        return [kw.pop(k) for k in keys]

    def longitudinal():
        if "espread" in kw or "blength" in kw:
            blength, espread = require('Both "blength" and "espread" are required',
                                       "blength", "espread")
            ems = blength * espread
            rs = (np.array([[espread / blength, 0.0], [0.0, blength / espread]])
                  if ems > 0.0 else np.zeros((2, 2)))
            return ems, rs
        return None, np.zeros((2, 2))

    def expand(rmat, rs):
        r6 = np.zeros((3, 6, 6))
        if rmat.shape[0] < 3:
            r6[:2, :4, :4] = rmat
            r6[2, 4:, 4:] = rs
        else:
            r6[:] = rmat
        return r6

    def from_modes(r6, emit, ems):
        if emit is None:
            emx, emy = require("Emittances must be provided", "emitx", "emity")
        else:
            emx = kw.pop("emitx", emit[0])
            emy = kw.pop("emity", emit[1])
            ems = emit[2] if ems is None else ems
        sig = emx * r6[0] + emy * r6[1]
        if ems is None:
            warn(AtWarning("Monochromatic beam: no energy spread"))
        else:
            sig = sig + ems * r6[2]
        return sig

    if ring is not None:
        kw["ring"] = ring
    if "beam" in kw:
        message = "'beam' is provided"
        b = np.asarray(kw.pop("beam"), dtype=np.float64)
        orbit = kw.pop("orbit", None)
        if orbit is None:
            orbit = np.mean(b, axis=1)
        d = b - np.asarray(orbit, dtype=np.float64).reshape(6, 1)
        sig = (d @ d.T) / d.shape[1]
    elif "twiss_in" in kw:
        message = "'twiss_in' is provided"
        tw = kw.pop("twiss_in")
        ems, rs = longitudinal()
        try:
            rin = np.asarray(tw["R"])
        except (ValueError, KeyError, IndexError):
            r6 = np.zeros((3, 6, 6))
            r6[0, :2, :2] = _twiss_block(tw["alpha"][0], tw["beta"][0])
            r6[1, 2:4, 2:4] = _twiss_block(tw["alpha"][1], tw["beta"][1])
            r6[2, 4:, 4:] = rs
        else:
            r6 = expand(rin, rs)
        sig = from_modes(r6, None, ems)
    elif "ring" in kw:
        from . import envelope, physics

        message = "'ring' is provided"
        ring = kw.pop("ring")
        orbit = kw.pop("orbit", None)
        ems, rs = longitudinal()
        if physics.is_6d(ring):
            _, beamdata, _ = envelope.ohmi_envelope(ring, orbit=orbit)
            r6 = expand(np.asarray(beamdata["mode_matrices"]), rs)
            sig = from_modes(r6, np.asarray(beamdata["mode_emittances"]), ems)
        else:
            m44, _ = physics.find_m44(ring, 0.0, orbit=orbit)
            a, _ = envelope.a_matrix(m44)
            r6 = expand(envelope.get_mode_matrices(a), rs)
            sig = from_modes(r6, None, ems)
    else:
        message = "Uncoupled beam"
        ax, ay, bx, by = require('Neither "twiss_in" nor "ring" is provided',
                                 "alphax", "alphay", "betax", "betay")
        ems, rs = longitudinal()
        r6 = np.zeros((3, 6, 6))
        r6[0, :2, :2] = _twiss_block(ax, bx)
        r6[1, 2:4, 2:4] = _twiss_block(ay, by)
        r6[2, 4:, 4:] = rs
        sig = from_modes(r6, None, ems)
    if kw:
        warn(AtWarning(f"{message}: {', '.join(kw)} ignored"))
    return sig


def _cholesky_planes(sigma: np.ndarray) -> np.ndarray:
    """Lower-triangular L with L L^T = sigma, falling back to the 4x4
    (x, y) or (x, delta) blocks and then to each plane when sigma is
    singular (a plane without emittance): particles.py:196-227."""

    def plane(slc):
        try:
            return cholesky(sigma[slc, slc])
        except LinAlgError:
            return np.zeros((2, 2))

    try:
        return cholesky(sigma)
    except LinAlgError:
        pass
    lmat = np.zeros((6, 6))
    for sel, rest in (([0, 1, 2, 3], slice(4, 6)), ([0, 1, 4, 5], slice(2, 4))):
        idx = np.ix_(sel, sel)
        try:
            lmat[idx] = cholesky(sigma[idx])
        except LinAlgError:
            lmat[:] = 0.0
            continue
        lmat[rest, rest] = plane(rest)
        return lmat
    for k in range(3):
        s = slice(2 * k, 2 * k + 2)
        lmat[s, s] = plane(s)
    return lmat


def beam(nparts: int, sigma, orbit=None, rng=None) -> np.ndarray:
    """(6, nparts) Fortran-ordered Gaussian particles with correlation
    matrix ``sigma``, centred on ``orbit`` (particles.py:181-234). ``rng``:
    a numpy Generator (default: a fresh one)."""
    sigma = np.asarray(sigma, dtype=np.float64)
    gen = rng if rng is not None else np.random.default_rng()
    v = gen.standard_normal((sigma.shape[1], int(nparts)))
    dist = np.asfortranarray(_cholesky_planes(sigma) @ v)
    if orbit is not None:
        dist += np.asarray(orbit, dtype=np.float64).reshape(6, 1)
    return dist


def emittances_from_beam(beam=None, sigma_mat=None):
    """Eigen-emittances and projected emittances (x, y, longitudinal) of a
    distribution or of a correlation matrix (particles.py:237-266)."""
    if beam is None:
        if sigma_mat is None:
            raise AtError("emittance calculation requires either beam or sigma_matrix")
    else:
        if sigma_mat is not None:
            warn(AtWarning("beam provided, sigma_matrix ignored in emittance calculation"))
        sigma_mat = sigma_matrix(beam=beam)
    sigma_mat = np.asarray(sigma_mat, dtype=np.float64)
    # the eigenvalues of Sigma J are +-i eps for the three modes: one of
    # each pair, the three largest imaginary parts (the reference's
    # np.unique of |eigenvalues| loses the pairing when a plane has no
    # emittance and rounding splits its pair; both agree otherwise)
    ev, _ = eig(sigma_mat @ _JMAT)
    eig_em = np.abs(np.sort(ev.imag)[-3:])
    blocks = [slice(0, 2), slice(2, 4), slice(6, 3, -1)]
    with np.errstate(invalid="ignore"):
        proj = np.sqrt([det(sigma_mat[s, s]) for s in blocks])
    # order the eigen-emittances like the projected ones (x, y, z)
    free = list(range(3))
    order = []
    for pk in np.nan_to_num(proj, nan=0.0):
        k = min(free, key=lambda j: abs(eig_em[j] - pk))
        order.append(k)
        free.remove(k)
    return eig_em[order], proj
