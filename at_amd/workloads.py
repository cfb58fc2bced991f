"""Synthetic workloads: the lattices and beams the engine is measured on.

``fodo_ring``: 32 FODO cells (320 elements), 3 GeV electrons, no radiation,
no cavity; ``large_ring``: 32 synthetic cells of about 62 elements each with
radiating bends and quadrupoles, sextupoles, octupoles, correctors, one
rectangular aperture per cell and one RF cavity; ``gaussian_beam``: seeded
6-D Gaussian particles in the (6, N) Fortran layout of pyat.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np

from .lattice import (Aperture, Corrector, Dipole, Drift, Lattice, Octupole,
                      Particle, Quadrupole, RFCavity, Sextupole, ThinMultipole)

SIGMA0 = np.array([1e-4, 1e-5, 1e-5, 1e-6, 1e-3, 1e-3])


def gaussian_beam(n: int, seed: int = 0, scale: float = 1.0) -> np.ndarray:
    """(6, n) F-ordered FP64 beam: r = (scale*sigma) * z, z ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((6, n))
    return np.asfortranarray((scale * SIGMA0)[:, None] * z)


_C_LIGHT = 299792458.0
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# Large ring: 32 copies of the frozen cell in data/large_ring_cell.json
# (derived from the ESRF-EBS HMBA cell by tests/golden/make_large_ring.py), radiating
# bends and quadrupoles, one RF cavity, one rectangular aperture per cell.
LARGE_RING = {
    "ncells": 32,
    "energy": 6.0e9,
    "harmonic_number": 992,
    "rf_frequency": 352.0e6,
    "rf_voltage": 6.0e6,
    # synchronous phase for U0 = 2.30 MeV/turn: c/(2 pi f) * asin(U0/V)
    "time_lag": 0.053337214233615114,
    "aperture": [-0.010, 0.010, -0.005, 0.005],
    "corrector_seed": 20241202,
    "corrector_rms": 1.0e-6,
    "nslices": 10,
}


def large_ring_cell() -> list[dict]:
    with open(os.path.join(_DATA, "large_ring_cell.json")) as f:
        return json.load(f)["elements"]


def large_ring(params: dict | None = None, radiation: bool = True, cavity: bool = True,
               ncells: int | None = None) -> Lattice:
    """About 2,240 elements: 32 cells of 7 bends (4 longitudinal-gradient, 3
    combined-function), 16 quadrupoles, 6 sextupoles, 2 octupoles, 3
    correctors, 1 rectangular aperture (+-10 mm x +-5 mm) and drifts; one RF
    cavity (352 MHz, h = 992, 6 MV) at the start; circumference h*c/f.
    With radiation the bends and quadrupoles use the Rad passes; sextupoles
    and octupoles never do (pyat lattice_object.py:53-79)."""
    p = dict(LARGE_RING)
    if params:
        p.update(params)
    nc = ncells or p["ncells"]
    ns = p["nslices"]
    bpass = "BndMPoleSymplectic4RadPass" if radiation else "BndMPoleSymplectic4Pass"
    qpass = "StrMPoleSymplectic4RadPass" if radiation else "StrMPoleSymplectic4Pass"
    rng = np.random.default_rng(p["corrector_seed"])
    cell = large_ring_cell()
    elems = []
    for _ in range(nc):
        elems.append(Aperture("AP", p["aperture"]))
        for it in cell:
            t = it["type"]
            if t == "drift":
                elems.append(Drift("D", it["L"]))
            elif t == "quad":
                elems.append(Quadrupole(it["name"], it["L"], it["k"], NumIntSteps=ns,
                                        PassMethod=qpass))
            elif t == "bend":
                elems.append(Dipole(it["name"], it["L"], it["angle"], it["k"],
                                    EntranceAngle=it["entrance_angle"],
                                    ExitAngle=it["exit_angle"], NumIntSteps=ns,
                                    PassMethod=bpass))
            elif t == "sext":
                elems.append(Sextupole(it["name"], it["L"], it["k2"], NumIntSteps=ns))
            elif t == "oct":
                elems.append(Octupole(it["name"], it["L"], [0, 0, 0, 0], [0, 0, 0, it["k3"]],
                                      NumIntSteps=ns))
            elif t == "corrector":
                elems.append(Corrector(it["name"], it["L"], rng.normal(0.0, p["corrector_rms"], 2)))
            else:
                raise ValueError(t)
    if cavity:
        elems.insert(0, RFCavity("RF", 0.0, p["rf_voltage"], p["rf_frequency"],
                                 p["harmonic_number"], p["energy"], TimeLag=p["time_lag"]))
    return Lattice(elems, energy=p["energy"], particle=Particle("electron"), name="large_ring")


def _truncnorm(rng, sigma: float, cut: float, size):
    z = rng.standard_normal(size)
    bad = np.abs(z) > cut
    while bad.any():
        z[bad] = rng.standard_normal(int(bad.sum()))
        bad = np.abs(z) > cut
    return sigma * z


def _roll(theta: float) -> np.ndarray:
    c, s = math.cos(theta), math.sin(theta)
    m = np.eye(6)
    m[0, 0] = m[1, 1] = m[2, 2] = m[3, 3] = c
    m[0, 2] = m[1, 3] = s
    m[2, 0] = m[3, 1] = -s
    return m


def error_realisation(ring: Lattice, seed: int, shift: float = 30e-6, roll: float = 100e-6,
                      gradient: float = 1e-3, cut: float = 2.5, settings=None) -> Lattice:
    """Copy of ``ring`` with seeded errors on every quadrupole and sextupole:
    transverse shifts (T1 = -d, T2 = +d), rolls about s (R1 = R(theta),
    R2 = R(-theta)) and relative strength errors on PolynomB[1]
    (quadrupoles) / PolynomB[2] (sextupoles), each Gaussian truncated at
    ``cut`` sigma. The draws do not depend on the sizes, so scaling all three
    scales one realisation. ``settings``: orbit-correction actuator settings
    (one per ``correction_actuators`` entry) applied to the copy."""
    rng = np.random.default_rng(seed)
    acts = correction_actuators(ring) if settings is not None else []
    own = {k for k, _ in acts}
    elems = []
    for i, e in enumerate(ring):
        if isinstance(e, (Quadrupole, Sextupole)):
            e = e.deepcopy()
            dx, dy = _truncnorm(rng, shift, cut, 2)
            th = float(_truncnorm(rng, roll, cut, 1)[0])
            dk = float(_truncnorm(rng, gradient, cut, 1)[0])
            e.T1 = np.array([-dx, 0.0, -dy, 0.0, 0.0, 0.0])
            e.T2 = np.array([dx, 0.0, dy, 0.0, 0.0, 0.0])
            e.R1 = _roll(th)
            e.R2 = _roll(-th)
            pb = e.PolynomB.copy()
            pb[2 if isinstance(e, Sextupole) else 1] *= 1.0 + dk
            e.PolynomB = pb
        elif i in own:
            e = e.deepcopy()
        elems.append(e)
    lat = ring.copy()
    list.__init__(lat, elems)
    if settings is not None:
        if len(settings) != len(acts):
            raise ValueError(f"{len(settings)} settings for {len(acts)} actuators")
        for (k, plane), v in zip(acts, settings):
            _set_kick(lat[k], plane, float(v))
    return lat


# --- Orbit correction of the error realisations --------------------------
# With 30 um rms shifts the large ring has a closed orbit of about 0.5-0.8 mm
# rms already at a tenth of the errors, and none at their full size: the
# realisations that are tracked for many turns are orbit-corrected first, as
# error studies of such rings are. Actuators: the horizontal and vertical
# kicks of the 3 correctors per cell and the dipole coils of the 6
# sextupoles per cell (PolynomB[0], PolynomA[0]); the orbit is read at the
# entrance of every quadrupole, sextupole and corrector (800 points). The
# errors are ramped up from zero with the correction applied at each step,
# (step halved when the closed orbit is lost), and the correction is the
# Tikhonov-regularised least-squares solution of the ideal ring's response
# matrix. The settings of 64 realisations are frozen in
# data/large_ring_corrections.npz (written by tests/golden/make_corrections.py
# on the engine).

def correction_actuators(ring) -> list:
    """(element index, plane) of every orbit-correction actuator, horizontal
    kicks first (0: x, 1: y): correctors, then sextupole dipole coils."""
    cors = [i for i, e in enumerate(ring) if isinstance(e, Corrector)]
    sexts = [i for i, e in enumerate(ring) if isinstance(e, Sextupole)]
    return [(k, p) for p in (0, 1) for k in cors + sexts]


def correction_bpms(ring) -> list:
    """Element indices where the orbit is read (their entrances)."""
    return [i for i, e in enumerate(ring) if isinstance(e, (Quadrupole, Sextupole, Corrector))]


def _kick(e, plane: int) -> float:
    """Kick angle of an actuator: KickAngle of a corrector; -L B0 (x) or
    +L A0 (y) of a sextupole's dipole coil (px -= L B0, py += L A0)."""
    if isinstance(e, Corrector):
        return float(e.KickAngle[plane])
    return float(-e.PolynomB[0] * e.Length if plane == 0 else e.PolynomA[0] * e.Length)


def _set_kick(e, plane: int, value: float) -> None:
    if isinstance(e, Corrector):
        ka = np.array(e.KickAngle, dtype=float)
        ka[plane] = value
        e.KickAngle = ka
    elif plane == 0:
        pb = np.array(e.PolynomB, dtype=float)
        pb[0] = -value / e.Length
        e.PolynomB = pb
    else:
        pa = np.array(e.PolynomA, dtype=float)
        pa[0] = value / e.Length
        e.PolynomA = pa


def _bpm_orbit(ring, bpms, guess):
    """Closed orbit (6,) and (x..., y...) at the bpms, or (None, None) when
    the 4-D closed-orbit search does not converge."""
    import warnings

    from . import physics

    with warnings.catch_warnings():
        warnings.simplefilter("error")
        try:
            o, b = physics.find_orbit4(ring, 0.0, bpms, guess=guess)
        except Exception:  # no closed orbit (non-convergence warning or a lost probe)
            return None, None
    if not np.all(np.isfinite(b)):
        return None, None
    return o, np.concatenate([b[:, 0], b[:, 2]])


def orbit_response(ring, kick: float = 1e-6) -> np.ndarray:
    """Response of the orbit at the bpms to each actuator, (2 nbpm,
    nactuator), by one-sided differences of ``kick`` on ``ring`` (4-D)."""
    bpms, acts = correction_bpms(ring), correction_actuators(ring)
    _, b0 = _bpm_orbit(ring, bpms, None)
    if b0 is None:
        raise ValueError("the ring has no closed orbit")
    elems = list(ring)
    out = np.empty((b0.size, len(acts)))
    for j, (k, plane) in enumerate(acts):
        e0 = elems[k]
        e1 = e0.deepcopy()
        _set_kick(e1, plane, _kick(e0, plane) + kick)
        ring[k] = e1
        try:
            _, b = _bpm_orbit(ring, bpms, None)
        finally:
            ring[k] = e0
        if b is None:
            raise ValueError(f"no closed orbit with actuator {j} kicked")
        out[:, j] = (b - b0) / kick
    return out


def correct_realisation(base, seed: int, response: np.ndarray, *, shift: float = 30e-6,
                        roll: float = 100e-6, gradient: float = 1e-3, reg: float = 1e-3,
                        step0: float = 0.03, min_step: float = 0.005, final: int = 8,
                        max_rms: float = 100e-6):
    """Actuator settings that correct the closed orbit of realisation
    ``seed`` of ``base`` (a 4-D ring: no radiation, no cavity), or None when
    the ramp fails. Each step raises the error scale by ``step`` (starting at
    ``step0``, x1.5 after a success up to 0.25, halved after a failure down
    to ``min_step``) and applies two corrections; at full size corrections
    continue while they reduce the orbit, at most ``final``. A realisation
    whose corrected orbit stays above ``max_rms`` rms in either plane counts
    as failed. Returns (settings, (rms x, rms y) of the corrected orbit at
    the bpms)."""
    bpms, acts = correction_bpms(base), correction_actuators(base)
    u, sv, vt = np.linalg.svd(response, full_matrices=False)
    gain = vt.T * (sv / (sv * sv + (sv[0] * reg) ** 2))

    def corrected(scale, settings, guess, iterations):
        ring = error_realisation(base, seed, shift * scale, roll * scale, gradient * scale,
                                 settings=settings)
        for _ in range(iterations):
            o, b = _bpm_orbit(ring, bpms, guess)
            if o is None:
                return None, None, None
            guess = o
            settings = settings - gain @ (u.T @ b)
            for (k, plane), v in zip(acts, settings):
                _set_kick(ring[k], plane, float(v))
        return settings, guess, ring

    settings = np.array([_kick(base[k], p) for k, p in acts])
    guess, done, step = np.zeros(6), 0.0, step0
    while done < 1.0:
        scale = min(1.0, done + step)
        s2, g2, _ = corrected(scale, settings, guess, 2)
        if s2 is None:
            step *= 0.5
            if step < min_step:
                return None
            continue
        settings, guess, done = s2, g2, scale
        step = min(0.25, 1.5 * step)
    n = len(bpms)

    def rms(settings, guess):
        ring = error_realisation(base, seed, shift, roll, gradient, settings=settings)
        o, b = _bpm_orbit(ring, bpms, guess)
        return (None, None) if o is None else ((float(np.std(b[:n])), float(np.std(b[n:]))), o)

    best, guess = rms(settings, guess)
    if best is None:
        return None
    for _ in range(final):
        s2, g2, _ = corrected(1.0, settings, guess, 1)
        if s2 is None:
            break
        r2, g3 = rms(s2, g2)
        if r2 is None or max(r2) >= max(best):
            break
        settings, guess, best = s2, g3, r2
    if max(best) > max_rms:
        return None
    return settings, best


_CORRECTIONS = os.path.join(_DATA, "large_ring_corrections.npz")


def frozen_corrections() -> dict:
    """The frozen orbit corrections of the large ring's error realisations:
    {"seeds": (n,), "settings": (n, nactuator), "orbit_rms": (n, 2), and the
    error sizes "shift", "roll", "gradient"}."""
    with np.load(_CORRECTIONS, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def corrected_realisations(ring: Lattice, n: int) -> list:
    """The first ``n`` frozen orbit-corrected error realisations of ``ring``
    (the large ring, with or without radiation and cavity)."""
    c = frozen_corrections()
    if n > len(c["seeds"]):
        raise ValueError(f"{len(c['seeds'])} corrected realisations are frozen, {n} asked")
    return [error_realisation(ring, int(c["seeds"][i]), float(c["shift"]), float(c["roll"]),
                              float(c["gradient"]), settings=c["settings"][i])
            for i in range(n)]


def da_polar_grid(n: int, xmax: float = 15e-3, ymax: float = 6e-3,
                  deltas=(-0.03, 0.0, 0.03)) -> np.ndarray:
    """(6, n) F-ordered dynamic-aperture grid: for each momentum deviation a
    radial grid r in [1e-6, 1] x theta in [0, pi], x = xmax r cos(theta),
    y = ymax r sin(theta) (the RADIAL mode of pyat acceptance/boundary.py:
    126-131), n points in total."""
    per = [n // len(deltas) + (1 if i < n % len(deltas) else 0) for i in range(len(deltas))]
    cols = []
    for d, m in zip(deltas, per):
        nth = max(2, int(round(math.sqrt(m))))
        nr = max(1, int(math.ceil(m / nth)))
        r = np.linspace(0.0, 1.0, nr)
        r[r == 0] = 1e-6
        th = np.linspace(0.0, math.pi, nth)
        g1, g2 = np.meshgrid(r, th)
        x = (xmax * np.cos(g2) * g1).ravel()[:m]
        y = (ymax * np.sin(g2) * g1).ravel()[:m]
        p = np.zeros((6, m))
        p[0] = x
        p[2] = y
        p[4] = d
        cols.append(p)
    return np.asfortranarray(np.concatenate(cols, axis=1))


def batched_da_workload(ninst: int = 64, seed: int = 7, nmin: int = 1000, nmax: int = 40000,
                        shift: float | None = None):
    """Error realisations of the large ring with DA grids of U[nmin, nmax]
    particles each: a list of (lattice, particles). Default: the frozen
    orbit-corrected realisations with 30 um shifts (corrected_realisations);
    ``shift``: uncorrected realisations with that rms shift instead (seeds
    1000 seed + i)."""
    rng = np.random.default_rng(seed)
    base = large_ring()
    if shift is None:
        lats = corrected_realisations(base, ninst)
    else:
        lats = [error_realisation(base, seed * 1000 + i, shift=shift) for i in range(ninst)]
    return [(lat, da_polar_grid(int(rng.integers(nmin, nmax + 1)))) for lat in lats]


def fodo_ring(ncells: int = 32, energy: float = 3.0e9) -> Lattice:
    """32 x [QF, D, B, D, QD, D, B, D, SF, SD], 6.0 m per cell."""
    angle = 2.0 * math.pi / (2 * ncells)
    elems = []
    for _ in range(ncells):
        elems += [
            Quadrupole("QF", 0.4, 1.0),
            Drift("D", 0.3),
            Dipole("B", 2.0, angle),
            Drift("D", 0.3),
            Quadrupole("QD", 0.4, -1.0),
            Drift("D", 0.3),
            Dipole("B", 2.0, angle),
            Drift("D", 0.3),
            ThinMultipole("SF", [0.0, 0.0, 0.0], [0.0, 0.0, 2.0]),
            ThinMultipole("SD", [0.0, 0.0, 0.0], [0.0, 0.0, -2.0]),
        ]
    return Lattice(elems, energy=energy, particle=Particle("electron"), name="fodo32")


def serving_trace(n_requests: int = 32, rate: float = 100.0, seed: int = 3, n_instances: int = 64,
                  nmin: int = 2000, nmax: int = 200000, turns=(256, 1024, 4096),
                  da_fraction: float = 0.5):
    """Request trace of a tracking service: Poisson arrivals at ``rate``
    per second, an error-realisation index in [0, n_instances), N
    log-uniform on [nmin, nmax], a turn count drawn from ``turns``; a
    fraction ``da_fraction`` are dynamic-aperture grids (da_polar_grid), the
    rest Gaussian beams. Returns a list of at_amd.serving.Request."""
    from .serving import Request

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        inst = int(rng.integers(0, n_instances))
        nt = int(rng.choice(np.asarray(turns)))
        if rng.uniform() < da_fraction:
            parts, kind = da_polar_grid(n), "da"
        else:
            parts, kind = gaussian_beam(n, seed=seed * 100003 + i), "track"
        out.append(Request(i, inst, parts, nt, arrival=t, kind=kind))
    return out
