"""Closed orbits and one-turn matrices by tracking on the engine.

Every evaluation of the one-turn map is one engine launch of a *probe*: the
point of interest plus, for each varied coordinate j, the two copies
``x +- step_j/2 e_j``. The probe gives at once the image of the point and
the central-difference Jacobian of the map there, and with refpts the same
launch gives the transfer matrices from the ring's entrance to every refpt.
Extra particles cost nothing on the device (a launch of 13 particles takes
as long as one of 1), so the closed-orbit solvers use the central
difference for their Newton Jacobian: it is second-order accurate where a
one-sided difference is first-order, and each Newton iteration, one launch,
gains more.

The closure conditions are written as one affine form,

    residual(x) = M(x)[rows] - L x - c,

solved for the coordinates ``free`` of the input:

* fixed momentum deviation (pyat ``find_orbit4`` with ``dp``,
  pyat/at/physics/orbit.py:16-72): rows = free = x, px, y, py; L = I;
* fixed path lengthening (``find_sync_orbit``, orbit.py:74-120): free =
  x..delta, rows = x, px, y, py, ct; L = I on the transverse rows and 0 on
  ct, c_ct = dct (the output ct itself equals dct);
* 6-D with cavities (``find_orbit6``, orbit.py:307-368): all six
  coordinates, L = I, c_ct = the ct lag of the harmonic number
  (beta c h / f_rf - L0).

Steps and defaults follow the reference options (``XYStep`` for the
transverse coordinates and for delta in the path-lengthening solve,
``DPStep`` for delta and ct in 6-D; ``convergence``, ``max_iterations``).
The matrices (matrix.py:19-205) are the probe's central differences with
the reference's steps, so they are the reference's own estimate.
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


# Tracking hooks: one turn in place (returns the refpt copies (6, N, R, 1)),
# one element, and the matrices of every element of a ring in one launch.
def _lpass(ring, r_in, refpts=None, keep_lattice=False):
    refs = np.zeros(0, np.uint32) if refpts is None else refpts
    return tracking.internal_lpass(ring, r_in, 1, refpts=refs, keep_lattice=keep_lattice)


def _epass(elem, r_in, **kwargs):
    return tracking.internal_epass(elem, r_in, **kwargs)


def _elements(ring, orbits, step, **kwargs):
    return tracking.element_matrices(ring, orbits, step=step, **kwargs)


class _Rows:
    """Rows [start, start + n) of a device buffer (N, 6), as track_batch takes
    particles (shape, dtype, ptr)."""

    def __init__(self, buf, start: int, n: int):
        self.shape = (n, 6)
        self.dtype = np.dtype(np.float64)
        self.ptr = buf.ptr + 48 * start


class _DeviceRings:
    """One-turn passes of the probes of many rings, one batched launch
    (atx_track_batch) per call; the device lattices are built once."""

    def __init__(self, rings, strict: bool = False):
        from . import engine

        kw = {}
        self.energy, self.rest, _ = tracking._energy_particle(rings[0], kw)
        for r in rings[1:]:
            e, rest, _ = tracking._energy_particle(r, {})
            if (e, rest) != (self.energy, self.rest):
                raise AtError("the rings of a batch must share energy and particle")
        self.strict = strict
        self.dls = [engine.DeviceLattice(list(r)) for r in rings]

    def turn(self, idx, probes) -> None:
        """Probes (6, k) of the rings ``idx`` through one turn, in place."""
        from . import device
        from .batch import track_batch

        k = probes[0].shape[1]
        host = np.concatenate([p.T for p in probes])
        buf = device.DeviceArray.from_host(host)
        try:
            track_batch([self.dls[i] for i in idx], [_Rows(buf, j * k, k) for j in range(len(idx))],
                        1, energy=self.energy, rest_energy=self.rest, losses=False,
                        strict=self.strict)
            out = buf.to_host()
        finally:
            buf.close()
        for j, p in enumerate(probes):
            p[:] = out[j * k:(j + 1) * k].T

    def along(self, i, orbit, refs) -> np.ndarray:
        r = np.asfortranarray(np.array(orbit, dtype=float).reshape(6, 1))
        rout = self.dls[i].track(r, 1, refs, energy=self.energy, rest_energy=self.rest,
                                 strict=self.strict)
        return rout[:, 0, :, 0].T.copy()

    def close(self) -> None:
        for dl in self.dls:
            dl.close()


# Tracking hook of the batched solvers: a factory of _DeviceRings-like objects.
_rings = _DeviceRings


def _refs(ring, refpts) -> np.ndarray:
    return tracking.get_uint32_index(ring, refpts)


_TRANSVERSE = np.array([1.0, 1.0, 1.0, 1.0, 0.0, 0.0])


def _probe_batch(center: np.ndarray, steps: np.ndarray):
    """Column 0 is ``center``; then, for the k coordinates with a nonzero
    step, k columns displaced by +step/2 and k by -step/2."""
    varied = np.flatnonzero(steps)
    k = varied.size
    batch = np.tile(np.asarray(center, dtype=float).reshape(6, 1), (1, 1 + 2 * k))
    half = 0.5 * steps[varied]
    batch[varied, 1 + np.arange(k)] += half
    batch[varied, 1 + k + np.arange(k)] -= half
    return np.asfortranarray(batch), varied


def _differences(images: np.ndarray, steps: np.ndarray, varied: np.ndarray) -> np.ndarray:
    """Central differences d image / d input[varied] from a probe's images
    (the leading axis may hold the six coordinates of any number of copies)."""
    k = varied.size
    return (images[..., 1:1 + k] - images[..., 1 + k:]) / steps[varied]


def _probe(ring, center, steps, refs=None, keep_lattice=False):
    """One launch: the image of ``center``, the Jacobian (6, k) of the map
    with respect to the varied coordinates, and the copies at ``refs``
    as (R, 6, 1 + 2k)."""
    batch, varied = _probe_batch(center, steps)
    copies = _lpass(ring, batch, refs, keep_lattice)
    copies = np.moveaxis(np.asarray(copies)[..., 0], 2, 0)
    return batch[:, 0].copy(), _differences(batch, steps, varied), copies


def _close(ring, x, free, rows, lin, const, steps, keep_lattice, conv, maxit):
    """Newton on residual(x) = M(x)[rows] - lin @ x - const over x[free]:
    one probe (one launch) per iteration. Stops when an update moves x by
    at most ``conv`` (2-norm); warns if ``maxit`` updates did not get there."""
    x = np.array(x, dtype=float)
    probe_steps = np.zeros(6)
    probe_steps[free] = steps[free]
    for _ in range(maxit):
        image, jac, _ = _probe(ring, x, probe_steps, None, keep_lattice)
        keep_lattice = True
        residual = image[rows] - lin @ x - const
        slope = jac[rows] - lin[:, free]
        update = np.linalg.solve(slope, residual)
        x[free] -= update
        if np.linalg.norm(update) <= conv:
            return x
    warnings.warn(AtWarning("Maximum number of iterations reached. Possible non-convergence"),
                  stacklevel=3)
    return x


def _options(kwargs, what):
    conv = kwargs.pop("convergence", DConstant.OrbConvergence)
    maxit = kwargs.pop("max_iterations", DConstant.OrbMaxIter)
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    dps = kwargs.pop("DPStep", DConstant.DPStep)
    method = kwargs.pop("method", ELossMethod.TRACKING) if what == "orbit6" else None
    if kwargs:
        raise AtError(f"Unexpected keywords for {what}: {', '.join(kwargs)}")
    return conv, maxit, xy, dps, method


def _start(guess) -> np.ndarray:
    return np.zeros(6) if guess is None else np.array(guess, dtype=float)


def _orbit_at_dp(ring, dp=None, guess=None, keep_lattice=False, **kwargs):
    """Transverse fixed point at a fixed momentum deviation."""
    conv, maxit, xy, _, _ = _options(kwargs, "orbit_dp")
    x = _start(guess)
    x[4] = 0.0 if dp is None else dp
    t = np.arange(4)
    lin = np.eye(6)[t]
    return _close(ring, x, t, t, lin, np.zeros(4), xy * np.ones(6), keep_lattice, conv, maxit)


def _orbit_at_dct(ring, dct=None, guess=None, keep_lattice=False, **kwargs):
    """Transverse fixed point whose one-turn path lengthening is ``dct``:
    delta is an unknown, the output ct must equal dct."""
    conv, maxit, xy, _, _ = _options(kwargs, "orbit_dct")
    x = _start(guess)
    rows = np.array([0, 1, 2, 3, 5])
    lin = np.zeros((5, 6))
    lin[np.arange(4), np.arange(4)] = 1.0
    const = np.zeros(5)
    const[4] = 0.0 if dct is None else dct
    return _close(ring, x, np.arange(5), rows, lin, const, xy * np.ones(6), keep_lattice, conv,
                  maxit)


def _orbit_6d(ring, cavpts=None, guess=None, keep_lattice=False, **kwargs):
    """6-D fixed point up to the ct lag of the RF harmonic number; the
    starting ct is the synchronous time lag from the energy loss."""
    conv, maxit, xy, dps, method = _options(kwargs, "orbit6")
    length = cell_length(ring)
    f_rf = get_rf_frequency(ring)
    harmonic = round(f_rf * length / beta(ring) / CLIGHT)
    if guess is None:
        x = np.zeros(6)
        try:
            _, lag = get_timelag_fromU0(ring, method=method, cavpts=cavpts)
        except AtError as exc:
            raise AtError("Could not determine the initial synchronous phase") from exc
        x[5] = -lag
    else:
        x = _start(guess)
    const = np.zeros(6)
    const[5] = beta(ring) * CLIGHT * harmonic / f_rf - length
    steps = xy * _TRANSVERSE + dps * (1.0 - _TRANSVERSE)
    every = np.arange(6)
    return _close(ring, x, every, every, np.eye(6), const, steps, keep_lattice, conv, maxit)


def _along(ring, orbit, refpts, keep_lattice):
    """The orbit at each refpt: one single-particle launch, (R, 6)."""
    refs = _refs(ring, refpts)
    if refs.size == 0:
        return np.empty((0, 6))
    r = np.asfortranarray(np.array(orbit, dtype=float).reshape(6, 1))
    copies = _lpass(ring, r, refs, keep_lattice)
    return np.asarray(copies)[:, 0, :, 0].T.copy()


def _momentum_choice(dp, dct, df):
    if sum(v is not None for v in (dp, dct, df)) > 1:
        raise AtError("For off-momentum specification, only one of dp, dct and df may be "
                      "specified")


def _dct_from_df(ring, df):
    f_rf = cell_revolution_frequency(ring) * cell_harmnumber(ring)
    return -cell_length(ring) * df / (f_rf + df)


def find_orbit4(ring, dp=None, refpts=None, *, dct=None, df=None, orbit=None,
                keep_lattice=False, **kwargs):
    """4-D closed orbit at a momentum deviation, or at a path lengthening
    (``dct``) or RF frequency shift (``df``) (orbit.py:123-212). Returns
    (orbit (6,), orbit at refpts (R, 6))."""
    _momentum_choice(dp, dct, df)
    if orbit is None:
        if df is not None:
            orbit = _orbit_at_dct(ring, _dct_from_df(ring, df), keep_lattice=keep_lattice,
                                  **kwargs)
        elif dct is not None:
            orbit = _orbit_at_dct(ring, dct, keep_lattice=keep_lattice, **kwargs)
        else:
            orbit = _orbit_at_dp(ring, dp, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _along(ring, orbit, refpts, keep_lattice)


def find_sync_orbit(ring, dct=None, refpts=None, *, dp=None, df=None, orbit=None,
                    keep_lattice=False, **kwargs):
    """Closed orbit at a path lengthening (orbit.py:215-304); ``dp`` or
    ``df`` may be given instead."""
    _momentum_choice(dp, dct, df)
    if orbit is None:
        if dp is not None:
            orbit = _orbit_at_dp(ring, dp, keep_lattice=keep_lattice, **kwargs)
        else:
            target = _dct_from_df(ring, df) if df is not None else dct
            orbit = _orbit_at_dct(ring, target, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _along(ring, orbit, refpts, keep_lattice)


def find_orbit6(ring, refpts=None, *, orbit=None, keep_lattice=False, **kwargs):
    """6-D closed orbit of a ring with cavities (orbit.py:371-458)."""
    if not _cavities(ring):
        raise AtError("No cavity found in the lattice")
    if orbit is None:
        orbit = _orbit_6d(ring, keep_lattice=keep_lattice, **kwargs)
        keep_lattice = True
    return orbit, _along(ring, orbit, refpts, keep_lattice)


def find_orbit(ring, refpts=None, **kwargs):
    """find_orbit6 on a 6-D lattice, find_sync_orbit when dct or df is
    given, find_orbit4 otherwise (orbit.py:461-519)."""
    if is_6d(ring):
        for k in ("dp", "dct", "df"):
            if kwargs.pop(k, None) is not None:
                raise AtError(f"{k} is not supported on a 6D lattice")
        return find_orbit6(ring, refpts=refpts, **kwargs)
    if kwargs.get("dct") is not None or kwargs.get("df") is not None:
        return find_sync_orbit(ring, refpts=refpts, **kwargs)
    return find_orbit4(ring, refpts=refpts, **kwargs)


# --- ring quantities ---------------------------------------------------------

def cell_length(ring) -> float:
    return float(sum(getattr(e, "Length", 0.0) for e in ring))


def beta(ring) -> float:
    """Relativistic beta, evaluated as pyat's Lattice.beta does
    (sqrt(1 - 1/gamma/gamma), lattice_object.py:911-919) so that the ct lag
    of the harmonic number is the reference's to the last bit."""
    rest = float(getattr(ring.particle, "rest_energy", 0.0))
    if rest == 0.0:
        return 1.0
    gamma = float(ring.energy / rest)
    return math.sqrt(1.0 - 1.0 / gamma / gamma)


def cell_revolution_frequency(ring) -> float:
    return beta(ring) * CLIGHT / cell_length(ring)


def _cavities(ring):
    return [e for e in ring if getattr(e, "PassMethod", "").endswith("CavityPass")]


def get_rf_frequency(ring) -> float:
    freqs = {float(c.Frequency) for c in _cavities(ring)}
    if not freqs:
        raise AtError("No cavity found in the lattice")
    if len(freqs) > 1:
        raise AtError("Cavities have different frequencies")
    return freqs.pop()


def cell_harmnumber(ring) -> int:
    return int(round(get_rf_frequency(ring) / cell_revolution_frequency(ring)))


def is_6d(ring) -> bool:
    """Cavities active or radiating passes present."""
    return any(getattr(e, "PassMethod", "").endswith(("RadPass", "CavityPass")) for e in ring)


def _signature(elem):
    """Hashable description of an element's pass and numeric attributes."""
    items = []
    for k, v in sorted(vars(elem).items()):
        if isinstance(v, np.ndarray):
            items.append((k, v.shape, v.tobytes()))
        elif isinstance(v, (int, float, str, bool, np.number)):
            items.append((k, v))
    return tuple(items)


def get_energy_loss(ring, method=ELossMethod.INTEGRAL) -> float:
    """Energy lost per turn [eV] (energy_loss.py:31-102).

    INTEGRAL: Cgamma E^4 / (2 pi) times the second radiation integral of
    the dipoles (angle^2 / length). TRACKING: the delta a particle starting
    at zero loses in each radiating element, summed; identical elements are
    passed once and counted as often as they occur."""
    periods = getattr(ring, "periodicity", 1)
    if method is ELossMethod.INTEGRAL:
        i2 = sum(e.BendingAngle ** 2 / e.Length for e in ring
                 if isinstance(e, Dipole) and e.Length > 0)
        return periods * CGAMMA / (2.0 * math.pi) * ring.energy ** 4 * i2
    if method is ELossMethod.TRACKING:
        counts, first = {}, {}
        for e in ring:
            if getattr(e, "PassMethod", "").endswith("RadPass"):
                key = _signature(e)
                counts[key] = counts.get(key, 0) + 1
                first.setdefault(key, e)
        lost = 0.0
        for key, e in first.items():
            r = np.zeros((6, 1), order="F")
            _epass(e, r, energy=ring.energy, particle=ring.particle)
            lost -= counts[key] * r[4, 0]
        return periods * lost * ring.energy
    raise AtError(f"Invalid method: {method}")


def get_timelag_fromU0(ring, *, method=ELossMethod.TRACKING, cavpts=None):
    """TimeLag of the cavities that puts the synchronous phase at zero, and
    its difference from the present TimeLag (energy_loss.py:105-206, for
    cavities of one frequency and one TimeLag: ct_s = c / (2 pi f_rf) *
    asin(U0 / V_rf))."""
    cav = _cavities(ring) if cavpts is None else [ring[i] for i in _refs(ring, cavpts)]
    freqs = {float(c.Frequency) for c in cav}
    lags = {float(getattr(c, "TimeLag", 0.0)) for c in cav}
    if len(freqs) > 1 or len(lags) > 1:
        raise AtError("cavities of several frequencies or time lags are not supported")
    periods = getattr(ring, "periodicity", 1)
    u0 = get_energy_loss(ring, method=method) / periods
    vrf = float(sum(c.Voltage for c in cav))
    if u0 > vrf:
        raise AtError(f"The RF voltage ({periods * vrf:.3e} eV) is lower than "
                      f"the radiation losses ({periods * u0:.3e} eV).")
    lag = CLIGHT / (2.0 * math.pi * freqs.pop()) * math.asin(u0 / vrf)
    return np.full(len(cav), lag), lag - lags.pop()


# --- transfer matrices --------------------------------------------------------

_J4 = np.kron(np.eye(2), np.array([[0.0, 1.0], [-1.0, 0.0]]))


def _symplectic_inverse(m: np.ndarray) -> np.ndarray:
    """Inverse of a symplectic 4x4 matrix, J^T m^T J."""
    return _J4.T @ m.T @ _J4


def find_m44(ring, dp=None, refpts=None, dct=None, df=None, orbit=None,
             keep_lattice=False, **kwargs):
    """4x4 one-turn matrix at the closed orbit (matrix.py:19-102), and the
    4x4 matrices from the entrance to each refpt; with ``full`` those are
    turned into one-turn matrices at the refpts (m M44 m^-1)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    full = kwargs.pop("full", False)
    if orbit is None:
        orbit, _ = find_orbit4(ring, dp=dp, dct=dct, df=df, keep_lattice=keep_lattice, XYStep=xy)
        keep_lattice = True
    steps = xy * _TRANSVERSE
    refs = _refs(ring, refpts)
    _, jac, copies = _probe(ring, orbit, steps, refs, keep_lattice)
    m44 = jac[:4]
    if refs.size == 0:
        return m44, np.empty((0, 4, 4))
    mstack = _differences(copies, steps, np.arange(4))[:, :4]
    if full:
        mstack = np.stack([m @ m44 @ _symplectic_inverse(m) for m in mstack])
    return m44, mstack


def find_m66(ring, refpts=None, orbit=None, keep_lattice=False, **kwargs):
    """6x6 one-turn matrix at the closed orbit and the matrices from the
    entrance to each refpt (matrix.py:105-168)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    dps = kwargs.pop("DPStep", DConstant.DPStep)
    if orbit is None:
        if is_6d(ring):
            orbit, _ = find_orbit6(ring, keep_lattice=keep_lattice, XYStep=xy, DPStep=dps,
                                   **kwargs)
        else:
            orbit, _ = find_orbit4(ring, keep_lattice=keep_lattice, XYStep=xy, **kwargs)
        keep_lattice = True
    steps = xy * _TRANSVERSE + dps * (1.0 - _TRANSVERSE)
    refs = _refs(ring, refpts)
    _, m66, copies = _probe(ring, orbit, steps, refs, keep_lattice)
    if refs.size == 0:
        return m66, np.empty((0, 6, 6))
    return m66, _differences(copies, steps, np.arange(6))


def find_elem_m66(elem, orbit=None, **kwargs):
    """6x6 matrix of one element around ``orbit`` (default zero) by central
    differences with step XYStep on every coordinate (matrix.py:171-205)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    steps = xy * np.ones(6)
    batch, varied = _probe_batch(np.zeros(6) if orbit is None else orbit, steps)
    _epass(elem, batch, **kwargs)
    return _differences(batch, steps, varied)


def element_matrices(ring, orbits=None, **kwargs):
    """find_elem_m66 of every element of ``ring`` at once: the 6x6 matrix of
    each element around the orbit at its entrance (``orbits`` (6, nelem),
    e.g. the closed orbit at every element; default 0) by central
    differences with step XYStep. pyat's _dmatr makes these calls element by
    element (pyat/at/physics/radiation.py:78-90, matrix.py:171-202); here
    they are one engine launch. Keywords: XYStep, energy, particle (default:
    the ring's), strict. Returns (nelem, 6, 6)."""
    xy = kwargs.pop("XYStep", DConstant.XYStep)
    return _elements(ring, orbits, xy, **kwargs)


def _close_batch(rings, xs, free, rows, lin, consts, steps, conv, maxit, strict):
    """_close for many rings at once: every Newton iteration is one batched
    launch of the probes of the rings not yet converged, each ring with its
    own state and stopping rule. Returns the tracker (the caller closes it)
    and the indices of the rings that did not converge."""
    probe_steps = np.zeros(6)
    probe_steps[free] = steps[free]
    tracker = _rings(rings, strict=strict)
    try:
        active = list(range(len(rings)))
        for _ in range(maxit):
            if not active:
                break
            batches = [_probe_batch(xs[i], probe_steps) for i in active]
            tracker.turn(active, [b for b, _ in batches])
            still = []
            for i, (probe, varied) in zip(active, batches):
                jac = _differences(probe, probe_steps, varied)
                residual = probe[rows, 0] - lin @ xs[i] - consts[i]
                update = np.linalg.solve(jac[rows] - lin[:, free], residual)
                xs[i, free] -= update
                if not np.linalg.norm(update) <= conv:
                    still.append(i)
            active = still
    except BaseException:
        tracker.close()
        raise
    return tracker, active


def _finish_batch(tracker, rings, xs, active, refpts):
    try:
        if active:
            warnings.warn(AtWarning(f"Maximum number of iterations reached for rings {active}. "
                                    "Possible non-convergence"), stacklevel=3)
        along = None
        if refpts is not None:
            along = [tracker.along(i, xs[i], _refs(rings[i], refpts)) for i in range(len(rings))]
    finally:
        tracker.close()
    return xs, along


def find_orbit4_batch(rings, dp=0.0, refpts=None, *, guesses=None, strict=False, **kwargs):
    """find_orbit4 at momentum deviation ``dp`` for many rings at once (the
    error realisations of a lattice, say): every Newton iteration is one
    batched launch of the probes of all rings not yet converged, each ring
    with its own Newton state, steps and stopping rule exactly as
    find_orbit4 (orbit.py:16-72). The rings must share energy and particle.
    ``guesses``: starting points (R, 6). Returns (orbits (R, 6), list of the
    orbits at ``refpts`` (nref, 6) per ring, or None without refpts); warns
    for the rings that did not converge (their last iterate is returned)."""
    conv, maxit, xy, _, _ = _options(kwargs, "orbit_dp")
    nr = len(rings)
    xs = np.zeros((nr, 6)) if guesses is None else np.array(guesses, dtype=float).reshape(nr, 6)
    xs[:, 4] = 0.0 if dp is None else dp
    t = np.arange(4)
    tracker, active = _close_batch(rings, xs, t, t, np.eye(6)[t], np.zeros((nr, 4)),
                                   xy * np.ones(6), conv, maxit, strict)
    return _finish_batch(tracker, rings, xs, active, refpts)


def find_orbit6_batch(rings, refpts=None, *, guesses=None, strict=False, **kwargs):
    """find_orbit6 for many rings with cavities at once (orbit.py:371-458),
    one batched launch per Newton iteration as find_orbit4_batch; each
    ring's starting ct is its own synchronous time lag unless ``guesses``
    are given. Returns (orbits (R, 6), orbits at refpts per ring or None)."""
    conv, maxit, xy, dps, method = _options(kwargs, "orbit6")
    nr = len(rings)
    xs = np.zeros((nr, 6))
    consts = np.zeros((nr, 6))
    for i, ring in enumerate(rings):
        if not _cavities(ring):
            raise AtError(f"No cavity found in the lattice of ring {i}")
        length = cell_length(ring)
        f_rf = get_rf_frequency(ring)
        harmonic = round(f_rf * length / beta(ring) / CLIGHT)
        consts[i, 5] = beta(ring) * CLIGHT * harmonic / f_rf - length
        if guesses is None:
            try:
                _, lag = get_timelag_fromU0(ring, method=method)
            except AtError as exc:
                raise AtError("Could not determine the initial synchronous phase") from exc
            xs[i, 5] = -lag
        else:
            xs[i] = np.asarray(guesses[i], dtype=float)
    steps = xy * _TRANSVERSE + dps * (1.0 - _TRANSVERSE)
    every = np.arange(6)
    tracker, active = _close_batch(rings, xs, every, every, np.eye(6), consts, steps, conv,
                                   maxit, strict)
    return _finish_batch(tracker, rings, xs, active, refpts)
