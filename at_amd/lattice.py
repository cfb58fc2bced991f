"""Element and lattice model: the parameter surface of the tracking path.

The classes mirror pyat's element constructors, attribute names, conversions
and defaults (pyat/at/lattice/elements/element_object.py:46-63,
basic_elements.py:240-445, magnet_elements.py:137-720) so that a pyat user's
lattice description carries over unchanged. Only the pass methods of the
tracking engine are modelled; the attribute values are what the engine's
element compiler reads (at_amd.engine).
"""

from __future__ import annotations

import copy as _copy
import warnings
from typing import Any, Iterable

import numpy as np
from scipy.constants import physical_constants as _cst

__all__ = [
    "AtError",
    "AtWarning",
    "Particle",
    "Element",
    "LongElement",
    "Marker",
    "Monitor",
    "Aperture",
    "Drift",
    "ThinMultipole",
    "Multipole",
    "Dipole",
    "Bend",
    "Quadrupole",
    "Sextupole",
    "Octupole",
    "RFCavity",
    "Corrector",
    "Lattice",
    "e_mass",
    "clight",
]

#: Electron rest energy [eV] (pyat/at/constants.py:9-10)
e_mass = 1.0e06 * _cst["electron mass energy equivalent in MeV"][0]
p_mass = 1.0e06 * _cst["proton mass energy equivalent in MeV"][0]
mu_mass = 1.0e06 * _cst["muon mass energy equivalent in MeV"][0]
clight = 2.99792458e8


class AtError(Exception):
    """Error raised by the tracking front-end."""


class AtWarning(UserWarning):
    """Warning raised by the tracking front-end."""


class Particle:
    """Circulating particle: rest energy [eV] and charge [e].

    Known names follow pyat/at/lattice/particle_object.py:13-21.
    """

    _known = {
        "relativistic": {"rest_energy": 0.0, "charge": -1.0},
        "electron": {"rest_energy": e_mass, "charge": -1.0},
        "positron": {"rest_energy": e_mass, "charge": 1.0},
        "proton": {"rest_energy": p_mass, "charge": 1.0},
        "posmuon": {"rest_energy": mu_mass, "charge": 1.0},
        "negmuon": {"rest_energy": mu_mass, "charge": -1.0},
    }

    def __init__(self, name: str | None = "relativistic", **kwargs):
        if name in self._known:
            kwargs.update(self._known[name])
        elif "rest_energy" not in kwargs or "charge" not in kwargs:
            raise AtError(f"Unknown particle {name!r}: give rest_energy and charge")
        self.name = name
        self.rest_energy = float(kwargs.pop("rest_energy"))
        self.charge = float(kwargs.pop("charge"))
        for k, v in kwargs.items():
            setattr(self, k, v)

    def __repr__(self) -> str:
        return f"Particle({self.name!r}, rest_energy={self.rest_energy}, charge={self.charge})"


def _array(value, shape=(-1,), dtype=np.float64):
    # F-ordered, aligned float64 as the C integrators read them
    # (pyat/at/lattice/elements/conversions.py:10-14)
    return np.require(value, dtype=dtype, requirements=["F", "A"]).reshape(shape, order="F")


def _array66(value):
    return _array(value, shape=(6, 6))


def _int(value, vmin: int | None = None):
    iv = int(value)
    if vmin is not None and iv < vmin:
        raise ValueError(f"Value must be greater of equal to {vmin}")
    return iv


def _nop(value):
    return value


class Element:
    """Base class of all elements: FamName, Length, PassMethod plus any
    keyword given, converted as pyat converts them."""

    _conversions: dict = {
        "FamName": str,
        "PassMethod": str,
        "Length": float,
        "R1": _array66,
        "R2": _array66,
        "T1": lambda v: _array(v, (6,)),
        "T2": lambda v: _array(v, (6,)),
        "RApertures": lambda v: _array(v, (4,)),
        "EApertures": lambda v: _array(v, (2,)),
        "KickAngle": lambda v: _array(v, (2,)),
        "PolynomB": _array,
        "PolynomA": _array,
        "BendingAngle": float,
        "MaxOrder": int,
        "NumIntSteps": lambda v: _int(v, vmin=0),
        "Energy": float,
        "EntranceAngle": float,
        "ExitAngle": float,
        "FullGap": float,
        "FringeInt1": float,
        "FringeInt2": float,
        "FringeBendEntrance": int,
        "FringeBendExit": int,
        "FringeQuadEntrance": int,
        "FringeQuadExit": int,
        "fringeIntM0": lambda v: _array(v, (5,)),
        "fringeIntP0": lambda v: _array(v, (5,)),
        "FieldScaling": float,
        "Voltage": float,
        "Frequency": float,
        "HarmNumber": int,
        "TimeLag": float,
        "PhaseLag": float,
        "Limits": lambda v: _array(v, (4,)),
    }

    def __init__(self, family_name: str, **kwargs):
        self.FamName = family_name
        self.Length = kwargs.pop("Length", 0.0)
        self.PassMethod = kwargs.pop("PassMethod", "IdentityPass")
        self.update(kwargs)

    def __setattr__(self, name: str, value: Any) -> None:
        try:
            value = self._conversions.get(name, _nop)(value)
        except Exception as exc:
            exc.args = (f"Element {getattr(self, 'FamName', '?')}, attribute {name}: {exc}",)
            raise
        super().__setattr__(name, value)

    def update(self, *args, **kwargs) -> None:
        for k, v in dict(*args, **kwargs).items():
            setattr(self, k, v)

    def copy(self) -> "Element":
        return _copy.copy(self)

    def deepcopy(self) -> "Element":
        return _copy.deepcopy(self)

    def items(self):
        return vars(self).items()

    def __repr__(self) -> str:
        attrs = ", ".join(f"{k}={v!r}" for k, v in vars(self).items() if k != "FamName")
        return f"{self.__class__.__name__}({self.FamName!r}, {attrs})"


class LongElement(Element):
    """Element with a length (pyat basic_elements.py:37-50)."""

    def __init__(self, family_name: str, length: float, *args, **kwargs):
        kwargs.setdefault("Length", length)
        super().__init__(family_name, *args, **kwargs)


class Marker(Element):
    """Zero-length marker, IdentityPass."""


class Monitor(Element):
    """Beam position monitor, IdentityPass."""


class Aperture(Element):
    """Rectangular transverse aperture [xmin, xmax, ymin, ymax]."""

    def __init__(self, family_name: str, limits, **kwargs):
        kwargs.setdefault("PassMethod", "AperturePass")
        super().__init__(family_name, Limits=limits, **kwargs)


class Drift(LongElement):
    """Field-free drift, DriftPass."""

    def __init__(self, family_name: str, length: float, **kwargs):
        kwargs.setdefault("PassMethod", "DriftPass")
        super().__init__(family_name, length, **kwargs)


class ThinMultipole(Element):
    """Thin multipole kick, ThinMPolePass.

    PolynomA/PolynomB are padded to a common length of
    max(MaxOrder, DefaultOrder, len-1)+1 as in pyat magnet_elements.py:236-255.
    """

    DefaultOrder = 0
    # ThinMPolePass reads BendingAngle as an array [bax, bay]
    # (ThinMPolePass.c:86-91).
    _conversions = dict(Element._conversions, BendingAngle=_array)

    def __init__(self, family_name: str, poly_a, poly_b, **kwargs):
        pa = _array(kwargs.pop("PolynomA", poly_a))
        pb = _array(kwargs.pop("PolynomB", poly_b))
        deforder = max(getattr(self, "DefaultOrder", 0), len(pa) - 1, len(pb) - 1)
        maxorder = int(kwargs.pop("MaxOrder", deforder))
        kwargs.setdefault("PassMethod", "ThinMPolePass")
        super().__init__(family_name, **kwargs)
        lab = max(maxorder, deforder) + 1
        self.MaxOrder = maxorder
        self.PolynomA = np.pad(pa, (0, lab - len(pa)))
        self.PolynomB = np.pad(pb, (0, lab - len(pb)))


class Multipole(LongElement, ThinMultipole):
    """Thick multipole, StrMPoleSymplectic4Pass with 10 slices by default."""

    _conversions = Element._conversions

    def __init__(self, family_name: str, length: float, poly_a, poly_b, **kwargs):
        kwargs.setdefault("PassMethod", "StrMPoleSymplectic4Pass")
        kwargs.setdefault("NumIntSteps", 10)
        kwargs.setdefault("Length", length)
        ThinMultipole.__init__(self, family_name, poly_a, poly_b, **kwargs)


class Dipole(Multipole):
    """Bending magnet, BndMPoleSymplectic4Pass (magnet_elements.py:464-562)."""

    DefaultOrder = 1

    def __init__(self, family_name: str, length: float, bending_angle: float = 0.0,
                 k: float = 0.0, **kwargs):
        kwargs.setdefault("BendingAngle", bending_angle)
        kwargs.setdefault("EntranceAngle", 0.0)
        kwargs.setdefault("ExitAngle", 0.0)
        kwargs.setdefault("PassMethod", "BndMPoleSymplectic4Pass")
        super().__init__(family_name, length, [], [0.0, k], **kwargs)

    @property
    def K(self) -> float:
        return float(self.PolynomB[1])


Bend = Dipole


class Quadrupole(Multipole):
    """Quadrupole, StrMPoleSymplectic4Pass (magnet_elements.py:596-645)."""

    DefaultOrder = 1

    def __init__(self, family_name: str, length: float, k: float = 0.0, **kwargs):
        super().__init__(family_name, length, [], [0.0, k], **kwargs)

    @property
    def K(self) -> float:
        return float(self.PolynomB[1])


class Sextupole(Multipole):
    """Sextupole, StrMPoleSymplectic4Pass (magnet_elements.py:648-675)."""

    DefaultOrder = 2

    def __init__(self, family_name: str, length: float, h: float = 0.0, **kwargs):
        super().__init__(family_name, length, [], [0.0, 0.0, h], **kwargs)

    @property
    def H(self) -> float:
        return float(self.PolynomB[2])


class Octupole(Multipole):
    """Octupole, StrMPoleSymplectic4Pass (magnet_elements.py:678-684)."""

    DefaultOrder = 3


class RFCavity(LongElement):
    """RF cavity, RFCavityPass (basic_elements.py:354-405)."""

    def __init__(self, family_name: str, length: float, voltage: float, frequency: float,
                 harmonic_number: int, energy: float, **kwargs):
        kwargs.setdefault("TimeLag", 0.0)
        kwargs.setdefault("PassMethod", "RFCavityPass")
        super().__init__(family_name, length, Voltage=voltage, Frequency=frequency,
                         HarmNumber=harmonic_number, Energy=energy, **kwargs)


class Corrector(LongElement):
    """Steering magnet, CorrectorPass (magnet_elements.py:686-710)."""

    def __init__(self, family_name: str, length: float, kick_angle, **kwargs):
        kwargs.setdefault("PassMethod", "CorrectorPass")
        super().__init__(family_name, length, KickAngle=kick_angle, **kwargs)


def _toggle_rad(pm: str, on: bool) -> str:
    if on:
        return pm if pm.endswith("RadPass") else pm[:-4] + "RadPass"
    return pm.replace("RadPass", "Pass")


class Lattice(list):
    """Sequence of elements with the ring parameters the tracking path reads:
    ``energy`` [eV], ``particle`` and ``periodicity``
    (pyat/at/lattice/lattice_object.py)."""

    def __init__(self, elems: Iterable[Element] = (), *, energy: float = 0.0,
                 particle: Particle | str = "relativistic", periodicity: int = 1,
                 name: str = "", **kwargs):
        super().__init__(elems)
        self.energy = float(energy)
        self.particle = particle if isinstance(particle, Particle) else Particle(particle)
        self.periodicity = int(periodicity)
        self.name = name
        for k, v in kwargs.items():
            setattr(self, k, v)

    @property
    def circumference(self) -> float:
        return self.periodicity * float(sum(getattr(e, "Length", 0.0) for e in self))

    @property
    def harmonic_number(self) -> int:
        cavs = [e for e in self if e.PassMethod.endswith("CavityPass")]
        return int(cavs[0].HarmNumber) * self.periodicity if cavs else 1

    def copy(self) -> "Lattice":
        return _copy.copy(self)

    def deepcopy(self) -> "Lattice":
        return _copy.deepcopy(self)

    def rotate(self, n: int) -> "Lattice":
        """New lattice starting at element n (lattice_object.py:537-542)."""
        if len(self) == 0:
            return self.copy()
        n = n % len(self)
        return self._with_elems(list(self[n:]) + list(self[:n]))

    def _with_elems(self, elems) -> "Lattice":
        lat = _copy.copy(self)
        list.__init__(lat, elems)
        return lat

    def radiation_on(self, copy: bool = False) -> "Lattice | None":
        """Switch dipoles and quadrupoles to their Rad passes and turn the
        cavities on (the "auto" mapping of lattice_object.py:53-79:
        sextupoles, octupoles and multipoles keep their pass)."""
        return self._switch(True, copy)

    def radiation_off(self, copy: bool = False) -> "Lattice | None":
        """Inverse of radiation_on: Rad passes off, cavities become
        IdentityPass (zero length) or DriftPass."""
        return self._switch(False, copy)

    enable_6d = radiation_on
    disable_6d = radiation_off

    def _switch(self, on: bool, copy: bool):
        elems = []
        for e in self:
            ne = e.deepcopy() if copy else e
            if isinstance(e, (Dipole, Quadrupole)) and e.PassMethod.endswith("Pass"):
                ne.PassMethod = _toggle_rad(e.PassMethod, on)
            elif isinstance(e, (Sextupole, Octupole, Multipole)) and not on:
                ne.PassMethod = _toggle_rad(e.PassMethod, False)
            elif isinstance(e, RFCavity):
                if on:
                    ne.PassMethod = "RFCavityPass"
                else:
                    ne.PassMethod = "IdentityPass" if e.Length == 0 else "DriftPass"
            elems.append(ne)
        if copy:
            return self._with_elems(elems)
        return None


def load_mat(filename: str, key: str | None = None, **kwargs) -> Lattice:
    """Read a MATLAB v5 lattice file (cell array of element structs) into a
    Lattice, mapping the 'Class' field onto the element classes above
    (pyat/at/load/matfile.py reads the same layout through scipy.io).

    Ring parameters (energy, periodicity) come from the element attributes
    ``Energy``/``Periodicity`` when a RingParam element is absent, as pyat
    does for files without one.
    """
    import scipy.io as sio

    mat = sio.loadmat(filename, squeeze_me=False, mat_dtype=True)
    if key is None:
        keys = [k for k in mat if not k.startswith("__")]
        key = "RING" if "RING" in keys else keys[0]
    cells = mat[key]
    classes = {"Drift": Drift, "Bend": Dipole, "Dipole": Dipole, "Quadrupole": Quadrupole,
               "Sextupole": Sextupole, "Octupole": Octupole, "Multipole": Multipole,
               "ThinMultipole": ThinMultipole, "Marker": Marker, "Monitor": Monitor,
               "RFCavity": RFCavity, "Corrector": Corrector, "Aperture": Aperture}
    elems = []
    energy = 0.0
    periodicity = 1
    for cell in cells.flat:
        st = cell[0, 0]
        attrs = {}
        for name in st.dtype.names:
            v = st[name]
            if v.dtype.kind in "UO":
                v = str(np.asarray(v).ravel()[0]) if v.size else ""
            else:
                v = np.asarray(v, dtype=np.float64)
                v = float(v.ravel()[0]) if v.size == 1 and name not in (
                    "PolynomA", "PolynomB", "KickAngle", "T1", "T2", "RApertures",
                    "EApertures", "Limits", "fringeIntM0", "fringeIntP0") else v.reshape(
                    v.shape, order="F")
            attrs[name] = v
        cls_name = attrs.pop("Class", None)
        fam = attrs.pop("FamName", "")
        if cls_name == "RingParam":
            # ring parameters, not an element (pyat/at/load/matfile.py:163-220)
            energy = float(attrs.get("Energy", energy))
            periodicity = int(attrs.get("Periodicity", periodicity))
            continue
        # K and H are views of PolynomB[1]/[2] in pyat, not separate data
        attrs.pop("K", None)
        attrs.pop("H", None)
        if "Energy" in attrs and energy == 0.0:
            energy = float(attrs["Energy"])
        if "Periodicity" in attrs:
            periodicity = int(attrs.pop("Periodicity"))
        for iname in ("MaxOrder", "NumIntSteps", "FringeBendEntrance", "FringeBendExit",
                      "FringeQuadEntrance", "FringeQuadExit", "HarmNumber"):
            if iname in attrs:
                attrs[iname] = int(attrs[iname])
        for aname in ("PolynomA", "PolynomB", "KickAngle", "T1", "T2", "RApertures",
                      "EApertures", "Limits"):
            if aname in attrs:
                attrs[aname] = np.asarray(attrs[aname], dtype=np.float64).ravel(order="F")
        for mname in ("R1", "R2"):
            if mname in attrs:
                attrs[mname] = np.asarray(attrs[mname], dtype=np.float64).reshape((6, 6), order="F")
        cls = classes.get(cls_name, Element)
        el = Element.__new__(cls)
        Element.__init__(el, fam, **attrs)
        elems.append(el)
    kwargs.setdefault("energy", energy)
    kwargs.setdefault("periodicity", periodicity)
    return Lattice(elems, **kwargs)


def _warn(msg: str) -> None:
    warnings.warn(AtWarning(msg), stacklevel=3)
