"""Derive the large-ring cell from the ESRF-EBS hybrid multi-bend achromat
cell (the reference's own test fixture machine_data/hmba.mat, kept as
tests/golden/hmba.mat) and freeze it as at_amd/data/large_ring_cell.json.

    python tests/golden/make_large_ring.py [path/to/hmba.mat]

Transformations (the cell keeps the HMBA optics, so it is stable and
chromatically corrected):
  * zero-length IdentityPass elements (markers, monitors, the cavity
    placeholder) are dropped and adjacent drifts merged;
  * the sliced bends become 7 bends (4 longitudinal-gradient dipoles
    DL1A/DL2B/DL2D/DL1E and 3 combined-function dipoles DQ1B/DQ2C/DQ1D):
    lengths and angles add, the gradient is the length-weighted mean, the
    entrance/exit edge angles are those of the outer slices;
  * the three zero-strength correction multipoles become CorrectorPass
    elements of the same length;
  * 10 integration slices for every thick magnet;
  * the two end drifts are lengthened equally so that 32 cells make
    h*c/f = 992*c/352 MHz = 844.864 m.
"""

from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(ROOT, "at_amd", "data", "large_ring_cell.json")

_GROUPS = [r"^DL1A_\d$", r"^DL2B_\d$", r"^DL2D_\d$", r"^DL1E_\d$", r"^(ml)?DQ1B$", r"^(ml)?DQ2C_\d$",
           r"^(ml)?DQ1D$"]


def _group(name: str):
    for i, pat in enumerate(_GROUPS):
        if re.match(pat, name):
            return i
    return None


def convert(hmba_path: str) -> list[dict]:
    sys.path.insert(0, ROOT)
    from at_amd.lattice import load_mat

    ring = load_mat(hmba_path)
    items: list[dict] = []
    for e in ring:
        pm = e.PassMethod
        if pm == "IdentityPass":
            continue
        if pm == "DriftPass":
            if items and items[-1]["type"] == "drift":
                items[-1]["L"] += float(e.Length)
            else:
                items.append({"type": "drift", "L": float(e.Length)})
            continue
        if pm.startswith("Bnd"):
            g = _group(e.FamName)
            if items and items[-1]["type"] == "bend" and items[-1]["group"] == g:
                b = items[-1]
                b["L"] += float(e.Length)
                b["angle"] += float(e.BendingAngle)
                b["kl"] += float(e.Length) * float(e.PolynomB[1])
                b["exit_angle"] = float(e.ExitAngle)
                b["fq_exit"] = int(getattr(e, "FringeQuadExit", 0))
            else:
                items.append({"type": "bend", "group": g, "name": re.sub(r"^ml|_\d$", "", e.FamName),
                              "L": float(e.Length), "angle": float(e.BendingAngle),
                              "kl": float(e.Length) * float(e.PolynomB[1]),
                              "entrance_angle": float(e.EntranceAngle),
                              "exit_angle": float(e.ExitAngle),
                              "fq_entrance": int(getattr(e, "FringeQuadEntrance", 0)),
                              "fq_exit": int(getattr(e, "FringeQuadExit", 0))})
            continue
        pb = np.asarray(e.PolynomB, dtype=float)
        order = int(e.MaxOrder)
        if not np.any(pb):
            items.append({"type": "corrector", "name": e.FamName, "L": float(e.Length)})
        elif order == 1:
            items.append({"type": "quad", "name": e.FamName, "L": float(e.Length), "k": float(pb[1]),
                          "fq_entrance": int(getattr(e, "FringeQuadEntrance", 0)),
                          "fq_exit": int(getattr(e, "FringeQuadExit", 0))})
        elif order == 2:
            items.append({"type": "sext", "name": e.FamName, "L": float(e.Length), "k2": float(pb[2])})
        else:
            items.append({"type": "oct", "name": e.FamName, "L": float(e.Length), "k3": float(pb[3])})
    for it in items:
        if it["type"] == "bend":
            it["k"] = it.pop("kl") / it["L"]
            del it["group"]
    return items


def main(argv):
    path = argv[1] if len(argv) > 1 else os.path.join(HERE, "hmba.mat")
    items = convert(path)
    lcell = 992 * 299792458.0 / 352.0e6 / 32
    total = sum(it["L"] for it in items)
    extra = lcell - total
    assert items[0]["type"] == "drift" and items[-1]["type"] == "drift"
    items[0]["L"] += extra / 2
    items[-1]["L"] += extra / 2
    with open(OUT, "w") as f:
        json.dump({"source": "machine_data/hmba.mat (ESRF-EBS HMBA cell)", "cell_length": lcell,
                   "elements": items}, f, indent=1)
    counts = {}
    for it in items:
        counts[it["type"]] = counts.get(it["type"], 0) + 1
    print(f"{len(items)} elements per cell, length {lcell:.6f} m (+{extra:.4f} m): {counts}")


if __name__ == "__main__":
    main(sys.argv)
