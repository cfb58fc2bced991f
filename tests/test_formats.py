"""CPU: lattice files. AT JSON and AT repr round trips (the reference's
own strategy, pyat/test/test_load_save.py:40-80) and files written in the
reference's layout (json.py:38-72, reprfile.py:77-100), checked by the
oracle's tracking of the loaded lattice."""

import os

import numpy as np
import pytest

from at_amd import formats, load_mat
from at_amd.lattice import AtError, Dipole, Quadrupole
from at_amd.workloads import error_realisation, fodo_ring, large_ring
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _lattices():
    return {
        "hmba": load_mat(os.path.join(GOLDEN, "hmba.mat")),
        "large_ring_errors": error_realisation(large_ring(ncells=2), 3),
        "fodo": fodo_ring(),
    }


def _same_elements(a, b):
    assert len(a) == len(b)
    for ea, eb in zip(a, b):
        assert type(ea) is type(eb), (ea.FamName, type(ea), type(eb))
        va = {k: v for k, v in vars(ea).items() if not k.startswith("_")}
        vb = {k: v for k, v in vars(eb).items() if not k.startswith("_")}
        assert set(va) == set(vb), (ea.FamName, set(va) ^ set(vb))
        for k in va:
            assert np.array_equal(np.asarray(va[k]), np.asarray(vb[k])), (ea.FamName, k)


@pytest.mark.parametrize("ext", [".json", ".repr"])
@pytest.mark.parametrize("name", ["hmba", "large_ring_errors", "fodo"])
def test_round_trip(tmp_path, ext, name):
    ring = _lattices()[name]
    fn = tmp_path / f"ring{ext}"
    formats.save_lattice(ring, fn)
    ring2 = formats.load_lattice(fn)
    _same_elements(ring, ring2)
    assert ring2.energy == ring.energy and ring2.periodicity == ring.periodicity
    assert ring2.particle.rest_energy == ring.particle.rest_energy
    p = np.asfortranarray(np.random.default_rng(1).normal(size=(6, 16)) * 1e-4)
    r1, r2 = p.copy(order="F"), p.copy(order="F")
    po.atpass(ring, r1, 2)
    po.atpass(ring2, r2, 2)
    np.testing.assert_array_equal(r1, r2)


REFERENCE_STYLE_JSON = """{
  "atjson": 1,
  "at_version": "0.6.1",
  "elements": [
    {"FamName": "D1", "Length": 0.5, "PassMethod": "DriftPass", "Class": "Drift"},
    {"FamName": "QF", "Length": 0.4, "PassMethod": "StrMPoleSymplectic4Pass",
     "MaxOrder": 1, "NumIntSteps": 10, "PolynomA": [0.0, 0.0], "PolynomB": [0.0, 1.2],
     "Class": "Quadrupole"},
    {"FamName": "B", "Length": 1.0, "PassMethod": "BndMPoleSymplectic4Pass",
     "BendingAngle": 0.05, "EntranceAngle": 0.025, "ExitAngle": 0.025, "MaxOrder": 1,
     "NumIntSteps": 10, "PolynomA": [0.0, 0.0], "PolynomB": [0.0, -0.2], "Class": "Bend"},
    {"FamName": "BPM", "Length": 0.0, "PassMethod": "IdentityPass", "Class": "Monitor"}
  ],
  "properties": {"name": "test", "energy": 3000000000.0, "periodicity": 1,
                 "particle": {"name": "electron", "rest_energy": 510998.95069, "charge": -1.0},
                 "harmonic_number": 400}
}
"""

REFERENCE_STYLE_REPR = """{'name': 'test', 'energy': 3000000000.0, 'periodicity': 1, 'particle': Particle('electron'), 'harmonic_number': 400}
Drift('D1', 0.5)
Quadrupole('QF', 0.4, 1.2, NumIntSteps=10)
Dipole('B', 1.0, 0.05, -0.2, EntranceAngle=0.025, ExitAngle=0.025)
Monitor('BPM')
"""


@pytest.mark.parametrize("ext", [".json", ".repr"])
def test_reference_layout(tmp_path, ext):
    fn = tmp_path / f"ring{ext}"
    fn.write_text(REFERENCE_STYLE_JSON if ext == ".json" else REFERENCE_STYLE_REPR)
    ring = formats.load_lattice(fn)
    assert [type(e).__name__ for e in ring] == ["Drift", "Quadrupole", "Dipole", "Monitor"]
    assert isinstance(ring[1], Quadrupole) and ring[1].K == 1.2
    assert isinstance(ring[2], Dipole) and ring[2].PolynomB[1] == -0.2
    assert ring.energy == 3e9 and ring.particle.rest_energy == pytest.approx(510998.95069)
    r = np.zeros((6, 1), order="F")
    r[0, 0] = 1e-3
    po.atpass(ring, r, 1, energy=ring.energy)
    assert np.all(np.isfinite(r))


def test_repr_is_parsed_without_eval(tmp_path):
    fn = tmp_path / "evil.repr"
    fn.write_text("{'energy': 1e9}\nDrift('D', __import__('os').getcwd())\n")
    with pytest.raises(AtError):
        formats.load_repr(fn)


def test_unknown_extension(tmp_path):
    with pytest.raises(AtError):
        formats.load_lattice(tmp_path / "ring.xyz")
