"""Element zoo: every in-scope pass method with its parameter variants.

Each case is (name, element, energy [eV]). The same cases drive the CPU
tests (oracle against the reference cores and the golden fixtures) and the
GPU tests (engine against the oracle).
"""

from __future__ import annotations

import math

import numpy as np

from at_amd.lattice import (Aperture, Corrector, Dipole, Drift, Element, Marker, Multipole,
                            Octupole, Quadrupole, RFCavity, Sextupole, ThinMultipole)

E3 = 3.0e9
E6 = 6.0e9


def _rot(theta: float) -> np.ndarray:
    """Roll about the beam axis as a 6x6 matrix (pyat transformation.py)."""
    c, s = math.cos(theta), math.sin(theta)
    m = np.eye(6)
    m[0, 0] = m[1, 1] = m[2, 2] = m[3, 3] = c
    m[0, 2] = m[1, 3] = s
    m[2, 0] = m[3, 1] = -s
    return m


def _misalign(e: Element, dx=3e-5, dy=-2e-5, roll=1e-4) -> Element:
    e.T1 = np.array([-dx, 0.0, -dy, 0.0, 0.0, 0.0])
    e.T2 = np.array([dx, 0.0, dy, 0.0, 0.0, 0.0])
    e.R1 = _rot(roll)
    e.R2 = _rot(-roll)
    return e


def _ap(e: Element) -> Element:
    e.RApertures = np.array([-2e-3, 2e-3, -1e-3, 1e-3])
    e.EApertures = np.array([3e-3, 1.5e-3])
    return e


def cases() -> list:
    c = []
    add = lambda name, e, energy=E3: c.append((name, e, energy))  # noqa: E731
    # --- IdentityPass / DriftPass / AperturePass
    add("marker", Marker("M"))
    add("identity_ap_tr", _ap(_misalign(Marker("M"))))
    add("drift", Drift("D", 1.3))
    add("drift_ap_tr", _ap(_misalign(Drift("D", 0.7))))
    add("aperture", Aperture("A", [-1e-3, 1.5e-3, -0.8e-3, 0.6e-3]))
    # --- StrMPoleSymplectic4Pass
    add("quad", Quadrupole("Q", 0.4, 1.2))
    add("quad_defocus", Quadrupole("Q", 0.5, -2.5, NumIntSteps=7))
    add("sext", Sextupole("S", 0.2, 80.0))
    add("oct", Octupole("O", 0.15, [0, 0, 0, 0.0], [0, 0, 0, 900.0]))
    add("mpole_order5_skew", Multipole("MP", 0.3, [0.0, 0.2, 3.0, 0.0, 0.0, 1e4],
                                       [0.001, 1.5, -20.0, 500.0, 0.0, -3e5]))
    add("quad_kick", Quadrupole("Q", 0.4, 1.0, KickAngle=np.array([2e-5, -1e-5])))
    add("quad_scaling", Quadrupole("Q", 0.4, 1.0, FieldScaling=1.002))
    add("quad_fringe_lee", Quadrupole("Q", 0.4, 2.0, FringeQuadEntrance=1, FringeQuadExit=1))
    add("quad_fringe_elegant", Quadrupole(
        "Q", 0.4, 2.0, FringeQuadEntrance=2, FringeQuadExit=2,
        fringeIntM0=np.array([0.01, 0.02, 0.003, 0.0004, 0.005]),
        fringeIntP0=np.array([0.011, -0.021, 0.0031, -0.00041, 0.0051])))
    add("quad_fringe_mixed", Quadrupole(
        "Q", 0.4, -1.5, FringeQuadEntrance=2, FringeQuadExit=1,
        fringeIntM0=np.array([0.01, 0.02, 0.003, 0.0004, 0.005]),
        fringeIntP0=np.array([0.011, -0.021, 0.0031, -0.00041, 0.0051])))
    add("quad_misaligned_ap", _ap(_misalign(Quadrupole("Q", 0.4, 1.7))))
    add("sext_misaligned", _misalign(Sextupole("S", 0.2, -60.0)))
    add("quad_nslice0", Quadrupole("Q", 0.4, 1.0, NumIntSteps=0))
    # --- BndMPoleSymplectic4Pass
    ang = 2 * math.pi / 64
    add("bend", Dipole("B", 2.0, ang))
    add("bend_cfgrad", Dipole("B", 1.2, 0.05, -0.6))
    add("bend_order0", Dipole("B", 0.5, 0.01, PolynomB=[0.0, 0.0], MaxOrder=0,
                              EntranceAngle=0.004, ExitAngle=0.002))
    add("bend_edges_m1", Dipole("B", 1.0, 0.1, EntranceAngle=0.05, ExitAngle=0.03))
    for m in (0, 1, 2, 3):
        add(f"bend_fringe_m{m}", Dipole("B", 1.0, 0.08, 0.3, EntranceAngle=0.04, ExitAngle=0.02,
                                        FullGap=0.03, FringeInt1=0.5, FringeInt2=0.4,
                                        FringeBendEntrance=m, FringeBendExit=m))
    add("bend_m2_nogap", Dipole("B", 1.0, 0.08, EntranceAngle=0.04, ExitAngle=0.02,
                                FringeBendEntrance=2, FringeBendExit=2))
    add("bend_quadfringe_kick", Dipole("B", 0.8, 0.03, -0.8, FringeQuadEntrance=1,
                                       FringeQuadExit=1, KickAngle=np.array([1e-5, 2e-5])))
    add("bend_misaligned_ap", _ap(_misalign(Dipole("B", 1.1, 0.04, 0.2))))
    add("bend_sext", Dipole("B", 1.0, 0.04, PolynomB=[0.0, -0.3, 5.0]))
    # --- radiating passes
    add("quad_rad", Quadrupole("Q", 0.4, 3.0, PassMethod="StrMPoleSymplectic4RadPass"), E6)
    add("sext_rad", Sextupole("S", 0.25, 120.0, PassMethod="StrMPoleSymplectic4RadPass"), E6)
    add("bend_rad", Dipole("B", 1.5, 0.06, -0.7, PassMethod="BndMPoleSymplectic4RadPass"), E6)
    add("bend_rad_edges", Dipole("B", 1.0, 0.05, -0.5, EntranceAngle=0.025, ExitAngle=0.025,
                                 PassMethod="BndMPoleSymplectic4RadPass"), E6)
    for m in (1, 2, 3):
        add(f"bend_rad_fringe_m{m}", Dipole("B", 1.0, 0.05, -0.5, EntranceAngle=0.025,
                                            ExitAngle=0.02, FullGap=0.025, FringeInt1=0.6,
                                            FringeInt2=0.5, FringeBendEntrance=m,
                                            FringeBendExit=m,
                                            PassMethod="BndMPoleSymplectic4RadPass"), E6)
    add("bend_rad_m2_nogap", Dipole("B", 1.0, 0.05, EntranceAngle=0.02, ExitAngle=0.01,
                                    FringeBendEntrance=2, FringeBendExit=2,
                                    PassMethod="BndMPoleSymplectic4RadPass"), E6)
    add("quad_rad_misaligned_scaled", _ap(_misalign(Quadrupole(
        "Q", 0.4, -2.2, FieldScaling=0.998, KickAngle=np.array([1e-5, 0.0]),
        FringeQuadEntrance=1, FringeQuadExit=1, PassMethod="StrMPoleSymplectic4RadPass"))), E6)
    # --- ThinMPolePass
    add("thin_sext", ThinMultipole("SF", [0, 0, 0], [0, 0, 2.0]))
    add("thin_kick_bend", ThinMultipole("T", [0, 1e-3, 0], [0, 0.01, -3.0],
                                        KickAngle=np.array([1e-4, -2e-4]),
                                        BendingAngle=np.array([1e-3, 5e-4])))
    add("thin_order0", ThinMultipole("T", [2e-5], [1e-4]))
    add("thin_scaled_misaligned", _misalign(ThinMultipole("T", [0, 0, 0, 10.0], [0, 0.1, 1.0, 50.0],
                                                          FieldScaling=1.01)))
    # --- RFCavityPass
    add("rf_thin", RFCavity("RF", 0.0, 187500.0, 3.5237e8, 31, 6.0e9), E6)
    add("rf_thick_lag", RFCavity("RF", 0.5, 6.0e6, 3.52e8, 992, 6.0e9, TimeLag=0.01,
                                 PhaseLag=0.1), E6)
    add("rf_zero_voltage_thick", RFCavity("RF", 0.4, 0.0, 3.52e8, 992, 6.0e9), E6)
    # --- CorrectorPass
    add("corrector_thin", Corrector("C", 0.0, np.array([1e-4, -2e-4])))
    add("corrector_thick", Corrector("C", 0.2, np.array([1e-4, -2e-4]), FieldScaling=1.001))
    add("corrector_thick_ap_tr", _ap(_misalign(Corrector("C", 0.1, np.array([3e-5, 1e-5])))))
    # --- CavityPass
    add("cavity_thin", RFCavity("C", 0.0, 2.0e6, 3.52e8, 992, 6.0e9, TimeLag=0.02,
                                PassMethod="CavityPass"), E6)
    add("cavity_thick_phase", RFCavity("C", 0.6, 4.0e6, 3.52e8, 992, 6.0e9, PhaseLag=0.2,
                                       PassMethod="CavityPass"), E6)
    # --- ExactDriftPass / ExactMultipolePass / ExactMultipoleRadPass
    add("exact_drift", Drift("D", 1.3, PassMethod="ExactDriftPass"))
    add("exact_drift_ap_tr", _ap(_misalign(Drift("D", 0.7, PassMethod="ExactDriftPass"))))
    add("exact_quad", Quadrupole("Q", 0.4, 1.2, PassMethod="ExactMultipolePass"))
    add("exact_quad_fringe", Quadrupole("Q", 0.5, -2.2, NumIntSteps=7, FringeQuadEntrance=1,
                                        FringeQuadExit=1, PassMethod="ExactMultipolePass"))
    add("exact_sext_kick_scaled", Sextupole("S", 0.2, 40.0, KickAngle=np.array([1e-4, -2e-4]),
                                            FieldScaling=1.01, PassMethod="ExactMultipolePass"))
    add("exact_mpole_skew_fringe", Multipole("M", 0.3, [0.0, 0.3, 4.0], [0.1, 0.9, -3.0, 50.0],
                                             FringeQuadEntrance=1,
                                             PassMethod="ExactMultipolePass"))
    add("exact_quad_misaligned_ap", _ap(_misalign(Quadrupole("Q", 0.4, 1.7,
                                                             PassMethod="ExactMultipolePass"))))
    add("exact_quad_rad", Quadrupole("Q", 0.4, 3.0, PassMethod="ExactMultipoleRadPass"), E6)
    add("exact_sext_rad_fringe", Sextupole("S", 0.25, 120.0, FringeQuadExit=1,
                                           PassMethod="ExactMultipoleRadPass"), E6)
    return c


def beam(n: int, seed: int, scale: float = 1.0, special: bool = True) -> np.ndarray:
    """(6, n) F-ordered test particles: Gaussian core, some at apertures,
    plus the special inputs at.c treats as lost (NaN, Inf, |r| > 1)."""
    rng = np.random.default_rng(seed)
    sig = scale * np.array([1e-3, 1e-4, 5e-4, 5e-5, 5e-3, 2e-3])
    r = sig[:, None] * rng.standard_normal((6, n))
    if special and n >= 8:
        r[0, 0] = np.nan
        r[5, 1] = np.inf
        r[1, 2] = 1.5
        r[4, 3] = -1.2
        r[:, 4] = 0.0
        r[0, 5] = 2e-3  # exactly on a rectangular limit
        r[2, 6] = -1e-3
        r[0, 7] = -np.inf
    return np.asfortranarray(r)
