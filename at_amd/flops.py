"""Algorithmic FP64 FLOP model of the tracking path.

Rules (counted per particle per element, independent of the implementation):
  * every FP64 + - * / is 1 FLOP (a*b+c = 2); sin/tan/sqrt/pow are not FLOPs;
  * anything that depends only on element parameters, ring constants or the
    turn number is precomputed: 0;
  * in non-radiating magnets 1/(1+delta) is formed once per element, in
    radiating ones once per change of delta;
  * the reference map sequence is kept: 4 drifts and 3 kicks per
    Forest-Ruth slice, full Horner evaluation to MaxOrder;
  * comparisons, loss checks and apertures are 0.
Per element with n = MaxOrder and s = NumIntSteps:
  DriftPass 14; StrMPoleSymplectic4Pass 6 + s(48 + 24n);
  BndMPoleSymplectic4Pass 14 + s(66 + 24n); StrMPoleSymplectic4RadPass
  2 + s(144 + 24n); BndMPoleSymplectic4RadPass 10 + s(177 + 24n);
  ThinMPolePass 8n + 2 (+8 with BendingAngle); RFCavityPass 32 (L > 0) or
  4 (L = 0); CavityPass 34 (L > 0) or 6 (L = 0); CorrectorPass 23 (thick)
  or 2 (thin); ExactDriftPass 14; ExactMultipolePass 1 + s(64 + 24n)
  (+ 54(n+1) + 30 per PTC fringe); ExactMultipoleRadPass 1 + s(161 + 24n)
  (+ fringes); Aperture/Identity 0; +6 per T1/T2, +66 per R1/R2.
These are the figures the benchmark divides by the kernel time to report
achieved FP64 FLOP/s; they are a property of the algorithm, not of the
kernel's instruction count.
"""

from __future__ import annotations

import numpy as np


def element_flops(e) -> int:
    pm = e.PassMethod
    n = int(getattr(e, "MaxOrder", 0))
    s = int(getattr(e, "NumIntSteps", 0))
    L = float(getattr(e, "Length", 0.0))
    if pm == "DriftPass":
        f = 14
    elif pm == "StrMPoleSymplectic4Pass":
        f = 6 + s * (48 + 24 * n)
    elif pm == "BndMPoleSymplectic4Pass":
        f = 14 + s * (66 + 24 * n)
    elif pm == "StrMPoleSymplectic4RadPass":
        f = 2 + s * (144 + 24 * n)
    elif pm == "BndMPoleSymplectic4RadPass":
        f = 10 + s * (177 + 24 * n)
    elif pm == "ThinMPolePass":
        f = 8 * n + 2 + (8 if hasattr(e, "BendingAngle") else 0)
    elif pm == "RFCavityPass":
        f = 32 if L > 0 else 4
    elif pm == "CorrectorPass":
        f = 23 if L != 0 else 2
    elif pm == "CavityPass":
        f = 34 if L > 0 else 6
    elif pm == "ExactDriftPass":
        f = 14
    elif pm in ("ExactMultipolePass", "ExactMultipoleRadPass"):
        f = 1 + s * ((161 if pm.endswith("RadPass") else 64) + 24 * n)
        nfringe = int(bool(getattr(e, "FringeQuadEntrance", 0))) + int(
            bool(getattr(e, "FringeQuadExit", 0)))
        f += nfringe * (54 * (n + 1) + 30)
    else:  # AperturePass, IdentityPass
        f = 0
    if pm not in ("AperturePass", "RFCavityPass", "CavityPass"):
        f += 6 * (getattr(e, "T1", None) is not None) + 6 * (getattr(e, "T2", None) is not None)
        f += 66 * (getattr(e, "R1", None) is not None) + 66 * (getattr(e, "R2", None) is not None)
    return f


def lattice_flops(lattice) -> np.ndarray:
    """FLOPs per element (array of len(lattice))."""
    return np.array([element_flops(e) for e in lattice], dtype=np.float64)


def tracked_flops_and_passes(lattice, nturns: int, turn0: int, lost, lost_turn, lost_elem,
                             start_elem: int = 0):
    """Algorithmic FLOPs and particle-element passes of one tracking call,
    counting each particle until (and including) the element where it was
    lost; particles lost before the call count nothing."""
    fe = lattice_flops(lattice)
    nelem = len(fe)
    cum = np.concatenate([[0.0], np.cumsum(fe)])
    per_turn = cum[-1]
    lost = np.asarray(lost, bool)
    lt = np.asarray(lost_turn, np.int64) - turn0
    le = np.asarray(lost_elem, np.int64)
    nsurv = int((~lost).sum())
    flops = nsurv * nturns * per_turn
    passes = nsurv * nturns * float(nelem)
    if lost.any():
        lt_l = lt[lost]
        le_l = le[lost]
        ok = lt_l >= 0
        flops += float(np.sum(lt_l[ok] * per_turn + cum[le_l[ok] + 1]))
        passes += float(np.sum(lt_l[ok] * nelem + le_l[ok] + 1))
    return flops, passes
