"""Synthetic workloads: the lattices and beams the engine is measured on.

``fodo_ring``: 32 FODO cells (320 elements), 3 GeV electrons, no radiation,
no cavity; ``large_ring``: 32 synthetic cells of about 62 elements each with
radiating bends and quadrupoles, sextupoles, octupoles, correctors, one
rectangular aperture per cell and one RF cavity; ``gaussian_beam``: seeded
6-D Gaussian particles in the (6, N) Fortran layout of pyat.
"""

from __future__ import annotations

import math

import numpy as np

from .lattice import (Aperture, Corrector, Dipole, Drift, Lattice, Multipole, Octupole,
                      Particle, Quadrupole, RFCavity, Sextupole, ThinMultipole)

SIGMA0 = np.array([1e-4, 1e-5, 1e-5, 1e-6, 1e-3, 1e-3])


def gaussian_beam(n: int, seed: int = 0, scale: float = 1.0) -> np.ndarray:
    """(6, n) F-ordered FP64 beam: r = (scale*sigma) * z, z ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((6, n))
    return np.asfortranarray((scale * SIGMA0)[:, None] * z)


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
