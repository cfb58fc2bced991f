"""Equilibrium beam envelope by Ohmi's formalism, a caller of the
per-element diffusion matrices (pyat/at/physics/radiation.py:77-276) and
the normal-mode analysis it reports with (pyat/at/physics/amat.py).

The ring is tracked once with every element as an observation point to get
the closed orbit at each entrance; the diffusion matrices of all elements
come from one device launch (tracking.diffusion_matrices) and the
element transfer matrices from central differences on the engine, both
through module hooks so the same code runs on other tracking. The
cumulative diffusion B_k = M_k B_{k-1} M_k^T + D_k and the Sylvester
equation R = M R M^T + B for the equilibrium moments are small dense
algebra on the host.
"""

from __future__ import annotations

import numpy as np
from scipy.linalg import block_diag, det, eig, inv, solve, solve_sylvester

from . import physics, tracking
from .lattice import AtError

__all__ = ["ENVELOPE_DTYPE", "a_matrix", "amat", "jmat", "jmatswap", "get_mode_matrices",
           "get_tunes_damp", "ohmi_envelope", "diffusion_matrices_along"]

# record layout of the envelope data (radiation.py:64-72)
ENVELOPE_DTYPE = [
    ("r66", np.float64, (6, 6)),
    ("r44", np.float64, (4, 4)),
    ("m66", np.float64, (6, 6)),
    ("orbit6", np.float64, (6,)),
    ("emitXY", np.float64, (2,)),
    ("emitXYZ", np.float64, (3,)),
]

_J2 = np.array([[0.0, 1.0], [-1.0, 0.0]])


def jmat(ind: int) -> np.ndarray:
    """Block-diagonal antisymmetric matrix with ind blocks [[0, 1], [-1, 0]]
    (amat.py:20-30)."""
    if ind not in (1, 2, 3):
        raise ValueError("jmat dimension must be 1, 2 or 3")
    return block_diag(*([_J2] * ind))


def jmatswap(ind: int) -> np.ndarray:
    """jmat with the longitudinal block transposed, for the (delta, ct)
    coordinate order (amat.py:33-37)."""
    blocks = [_J2] * ind
    if ind == 3:
        blocks[2] = _J2.T
    return block_diag(*blocks)


def _plane_probes(dms: int) -> np.ndarray:
    """Column p is the complex vector -i e_{2p} + e_{2p+1}: the projection
    of a normalised eigenvector on plane p (amat.py:10,102)."""
    w = np.zeros((2 * dms, dms), dtype=complex)
    for p in range(dms):
        w[2 * p, p] = -1j
        w[2 * p + 1, p] = 1.0
    return w


def a_matrix(M):
    """A matrix and eigenvalues of a one-turn map, M = A R A^-1 with R a
    block diagonal of (damped) rotations (amat.py:40-119, Forest, Phys.
    Rev. E 58, 2481).

    Each complex-conjugate pair of eigenvectors is represented by the one
    of positive symplectic norm n = -i/2 v^H S v, normalised to |n| = 1;
    the modes are then assigned greedily to x, y, z by their largest
    projection, and A holds the real and imaginary parts of each mode's
    vector as consecutive columns."""
    M = np.asarray(M, dtype=np.float64)
    nv = M.shape[0]
    dms = nv // 2
    S = jmatswap(dms)
    lam, vec = eig(M)
    norm = (-0.5j * np.einsum("ik,ij,jk->k", vec.conj(), S, vec)).real
    if np.any(np.abs(norm) < 1.0e-12):
        raise AtError("Unstable ring")
    pos = np.flatnonzero(norm > 0.0)
    if pos.size != dms:
        raise AtError("Unstable ring")
    vn = vec[:, pos] / np.sqrt(norm[pos])
    lam = lam[pos]
    proj = np.abs(vn.conj().T @ S @ _plane_probes(dms))
    free = list(range(dms))
    chosen = []
    for p in range(dms):
        k = free[int(np.argmax(proj[free, p]))]
        chosen.append(k)
        free.remove(k)
    A = np.empty((nv, nv))
    A[:, 0::2] = vn[:, chosen].real
    A[:, 1::2] = vn[:, chosen].imag
    return A, lam[chosen]


def amat(M) -> np.ndarray:
    """A matrix only (amat.py:122-134)."""
    return a_matrix(M)[0]


def get_mode_matrices(A) -> np.ndarray:
    """R-matrices of the normal modes, R_k = A S I_k A^-1 S^T, valid for a
    non-symplectic A too (amat.py:163-189, Wolski, PRST-AB 9, 024001)."""
    A = np.asarray(A)
    dms = A.shape[0] // 2
    S = jmat(dms)
    T = jmatswap(dms)
    ainv_st = solve(A, S.T)
    modes = []
    for k in range(dms):
        sl = slice(2 * k, 2 * k + 2)
        modes.append((A[:, sl] @ T[sl, sl]) @ ainv_st[sl, :])
    return np.stack(modes, axis=0)


def get_tunes_damp(M, R=None):
    """Tunes, damping rates, mode matrices and (with a beam matrix R) mode
    emittances of a one-turn map (amat.py:198-247)."""
    nv = M.shape[0]
    dms = nv // 2
    A, lam = a_matrix(M)
    tunes = np.mod(np.angle(lam) / 2.0 / np.pi, 1.0)
    damping = -np.log(np.abs(lam))
    fields = [("tunes", np.float64, (dms,)), ("damping_rates", np.float64, (dms,)),
              ("mode_matrices", np.float64, (dms, nv, nv))]
    values = [tunes, damping, get_mode_matrices(A)]
    if R is not None:
        S = jmat(dms)
        rdiag = np.diag(A.T @ S @ R @ S @ A)
        fields.append(("mode_emittances", np.float64, (dms,)))
        values.append(-0.5 * (rdiag[0::2] + rdiag[1::2]))
    return np.rec.fromarrays(values, dtype=fields)


# --- hooks: the engine by default -------------------------------------------

def _diffmats(ring, orbits):
    """Diffusion matrices of every element, orbits (6, nelem). As in
    radiation.py:98-115 the energy is passed without a particle."""
    return tracking.diffusion_matrices(ring, orbits, energy=ring.energy, keep_lattice=True)


def _elem_m66s(ring, orbits, xystep):
    """Central-difference 6x6 matrix of every element from its entrance
    orbit (find_elem_m66, matrix.py:171-205), on one device lattice."""
    from . import engine

    part = getattr(ring, "particle", None)
    rest = float(getattr(part, "rest_energy", 0.0)) if part is not None else 0.0
    dl = engine.DeviceLattice(list(ring))
    try:
        scaling = xystep * np.ones(6)
        dg = 0.5 * np.diag(scaling)
        dmat = np.concatenate((dg, -dg), axis=1)
        out = np.empty((len(ring), 6, 6))
        for i in range(len(ring)):
            r = np.asfortranarray(orbits[:, i].reshape(6, 1) + dmat)
            dl.elempass(i, r, energy=float(ring.energy), rest_energy=rest)
            out[i] = (r[:, :6] - r[:, 6:]) / scaling
        return out
    finally:
        dl.close()


def diffusion_matrices_along(ring, orbit=None, keep_lattice: bool = False):
    """Cumulative diffusion matrices from the start of the ring to each
    element entrance and the closed orbit there (radiation.py:77-125):
    returns bbcum (nelem+1, 6, 6), with bbcum[0] = 0, and orbs (nelem+1, 6)."""
    if orbit is None:
        orbit, _ = physics.find_orbit6(ring, keep_lattice=keep_lattice)
        keep_lattice = True
    n = len(ring)
    refs = np.arange(n + 1, dtype=np.uint32)
    rout = physics._lpass(ring, np.asfortranarray(np.array(orbit, dtype=np.float64).reshape(6, 1)),
                          refs, keep_lattice)
    orbs = np.squeeze(rout, axis=(1, 3)).T
    entry = np.asfortranarray(orbs[:n].T)
    bb = _diffmats(ring, entry)
    ms = _elem_m66s(ring, entry, physics.DConstant.XYStep)
    bbcum = np.empty((n + 1, 6, 6))
    bbcum[0] = 0.0
    cum = bbcum[0]
    for i in range(n):
        cum = ms[i].dot(cum).dot(ms[i].T) + bb[i]
        bbcum[i + 1] = cum
    return bbcum, orbs


_SUBMAT = [slice(0, 2), slice(2, 4), slice(6, 3, -1)]


def _process(r66):
    """Projected emittances and the betatron (dp = 0) matrix of a beam
    matrix (radiation.py:210-230)."""
    emit3 = np.sqrt(np.maximum(np.array([det(r66[s, s]) for s in _SUBMAT]), 0.0))
    if emit3[0] < 1.0e-13:
        r44 = np.full((4, 4), np.nan)
    elif emit3[1] < 1.0e-13:
        minv = inv(r66[[0, 1, 4, 5], :][:, [0, 1, 4, 5]])
        r44 = np.zeros((4, 4))
        r44[:2, :2] = inv(minv[:2, :2])
    else:
        minv = inv(r66)
        r44 = inv(minv[:4, :4])
    emit2 = np.sqrt(np.maximum(np.array([det(r44[s, s], check_finite=False)
                                         for s in _SUBMAT[:2]]), 0.0))
    return r44, emit2, emit3


def ohmi_envelope(ring, refpts=None, orbit=None, keep_lattice: bool = False):
    """Equilibrium beam envelope (radiation.py:150-276; K. Ohmi et al.,
    Phys. Rev. E 49 (1994)): returns (emit0, beamdata, emit) as record
    arrays with the reference's fields. The ring must be 6D."""
    if not physics.is_6d(ring):
        raise AtError("ohmi_envelope needs radiation ON")
    refs = tracking.get_uint32_index(ring, refpts)
    bbcum, orbs = diffusion_matrices_along(ring, orbit=orbit, keep_lattice=keep_lattice)
    mring, ms = physics.find_m66(ring, refs, orbit=orbs[0], keep_lattice=True)
    # R = M R M^T + B as the Sylvester equation M^-1 R - R M^T = M^-1 B
    aa = inv(mring)
    rr = solve_sylvester(aa, -mring.T, aa @ bbcum[-1])
    rr = 0.5 * (rr + rr.T)
    rr4, emitxy, emitxyz = _process(rr)
    beamdata = get_tunes_damp(mring, rr)
    emit0 = np.rec.fromarrays((rr, rr4, mring, orbs[0], emitxy, emitxyz), dtype=ENVELOPE_DTYPE)
    if refs.size == 0:
        emit = np.recarray((0,), dtype=ENVELOPE_DTYPE)
    else:
        recs = []
        for m, cumb, orb in zip(ms, bbcum[refs], orbs[refs, :]):
            sig = m.dot(rr).dot(m.T) + cumb
            r44, e2, e3 = _process(sig)
            recs.append((sig, r44, m, orb, e2, e3))
        emit = np.rec.fromrecords(recs, dtype=ENVELOPE_DTYPE)
    return emit0, beamdata, emit
