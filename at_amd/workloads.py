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

from .lattice import (Aperture, Corrector, Dipole, Drift, Lattice, Multipole, Octupole,
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
# (derived from the ESRF-EBS HMBA cell by data/make_large_ring.py), radiating
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
