"""Route the orbit, matrix and acceptance drivers through the oracle
(test infrastructure): the CPU tests check the drivers' logic against the
reference's known answers with the oracle doing the tracking; the GPU tests
run the same drivers on the engine."""

import numpy as np

from at_amd import acceptance, physics
from oracle import pyoracle as po


def _ep(lattice):
    e = float(getattr(lattice, "energy", 0.0))
    return e, getattr(lattice, "particle", None)


def oracle_lpass(ring, r_in, refpts=None, keep_lattice=False):
    refs = np.zeros(0, np.uint32) if refpts is None else np.asarray(refpts, np.uint32)
    e, part = _ep(ring)
    return po.atpass(ring, r_in, 1, refpts=refs, energy=e, particle=part)


def oracle_epass(elem, r_in, energy=None, particle=None, **kwargs):
    return po.elempass(elem, r_in, energy=energy or 0.0, particle=particle)


def oracle_survived(parts, ring, nturns, **kwargs):
    e, part = _ep(ring)
    _, lm = po.atpass(ring, parts, nturns, losses=True, energy=e, particle=part, nthreads=8)
    return ~lm["islost"]


def oracle_project(ring, parts, obs, **kwargs):
    e, part = _ep(ring)
    po.atpass(ring, parts, 1, start_elem=int(obs), energy=e, particle=part)


class OracleRings:
    """The batched solvers' tracking hook on the oracle: each ring's probes
    through one turn, ring by ring."""

    def __init__(self, rings, strict=False):
        self.rings = rings

    def turn(self, idx, probes):
        for i, p in zip(idx, probes):
            e, part = _ep(self.rings[i])
            po.atpass(self.rings[i], p, 1, energy=e, particle=part)

    def along(self, i, orbit, refs):
        e, part = _ep(self.rings[i])
        r = np.asfortranarray(np.array(orbit, dtype=float).reshape(6, 1))
        rout = po.atpass(self.rings[i], r, 1, refpts=np.asarray(refs, np.uint32), energy=e,
                         particle=part)
        return np.asarray(rout)[:, 0, :, 0].T.copy()

    def close(self):
        pass


def install(monkeypatch):
    monkeypatch.setattr(physics, "_lpass", oracle_lpass)
    monkeypatch.setattr(physics, "_epass", oracle_epass)
    monkeypatch.setattr(physics, "_rings", OracleRings)
    monkeypatch.setattr(acceptance, "_survived", oracle_survived)
    monkeypatch.setattr(acceptance, "_project", oracle_project)
