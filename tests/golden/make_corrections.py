"""Orbit-correct the error realisations of the large ring on the engine and
freeze the actuator settings as at_amd/data/large_ring_corrections.npz.

    python tests/golden/make_corrections.py [count] [first seed] [out.npz]

Realisations: quadrupole and sextupole shifts 30 um rms, rolls 100 urad,
relative gradient errors 1e-3 (Gaussian, cut at 2.5 sigma), seeds from
`first seed` (default 7000) upwards; a seed whose ramp loses the closed
orbit, or whose corrected orbit stays above 100 um rms, is skipped, until
`count` (default 64) are corrected. The correction
(at_amd.workloads.correct_realisation) runs on the 4-D ring (no radiation,
no cavity) with the response matrix of the ideal ring; every closed-orbit
search is a GPU launch, so this needs a GPU (about 5 minutes on one MI355X).
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from at_amd import workloads  # noqa: E402


def main() -> None:
    count = int(sys.argv[1]) if len(sys.argv) > 1 else 64
    seed = int(sys.argv[2]) if len(sys.argv) > 2 else 7000
    out = sys.argv[3] if len(sys.argv) > 3 else os.path.join(
        os.path.dirname(os.path.dirname(HERE)), "at_amd", "data", "large_ring_corrections.npz")
    base = workloads.large_ring(radiation=False, cavity=False)
    t0 = time.time()
    response = workloads.orbit_response(base)
    print(f"response matrix {response.shape} in {time.time() - t0:.1f} s", flush=True)
    seeds, settings, rms, skipped = [], [], [], []
    while len(seeds) < count:
        t = time.time()
        res = workloads.correct_realisation(base, seed, response)
        if res is None:
            skipped.append(seed)
            print(f"seed {seed}: not corrected, skipped ({time.time() - t:.1f} s)",
                  flush=True)
        else:
            seeds.append(seed)
            settings.append(res[0])
            rms.append(res[1])
            print(f"seed {seed}: orbit rms x {res[1][0] * 1e6:.1f} um, y {res[1][1] * 1e6:.1f} um, "
                  f"settings rms {np.std(res[0]) * 1e6:.1f} urad ({time.time() - t:.1f} s)",
                  flush=True)
        seed += 1
    np.savez(out, seeds=np.array(seeds, np.int64), settings=np.array(settings),
             orbit_rms=np.array(rms), skipped=np.array(skipped, np.int64),
             shift=np.float64(30e-6), roll=np.float64(100e-6), gradient=np.float64(1e-3))
    print(f"{len(seeds)} realisations ({len(skipped)} skipped) -> {out} in "
          f"{time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
