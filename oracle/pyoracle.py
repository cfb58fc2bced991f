"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the CPU oracle.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module; it is the checker, never the product path.

``atpass``/``elempass`` take the same element objects and keyword meaning as
pyat's ``at.tracking.atpass.atpass``/``elempass`` (pyat/at.c:340-345,683)
and run either
  impl="oracle": the C restatement in oracle/oracle.c (liboracle.so), or
  impl="ref":    the reference integrator cores compiled from
                 /root/reference into oracle/_ref/ (ref_driver.c).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

PASS_IDS = {
    "IdentityPass": 0,
    "DriftPass": 1,
    "BndMPoleSymplectic4Pass": 2,
    "BndMPoleSymplectic4RadPass": 3,
    "StrMPoleSymplectic4Pass": 4,
    "StrMPoleSymplectic4RadPass": 5,
    "CavityPass": 6,
    "RFCavityPass": 7,
    "ExactDriftPass": 8,
    "ExactMultipolePass": 9,
    "ExactMultipoleRadPass": 10,
    "ThinMPolePass": 11,
    "CorrectorPass": 12,
    "AperturePass": 13,
}

_dp = C.POINTER(C.c_double)


class OracleElem(C.Structure):
    _fields_ = [
        ("pass_", C.c_int32), ("MaxOrder", C.c_int32), ("NumIntSteps", C.c_int32),
        ("FringeBendEntrance", C.c_int32), ("FringeBendExit", C.c_int32),
        ("FringeQuadEntrance", C.c_int32), ("FringeQuadExit", C.c_int32), ("pad0", C.c_int32),
        ("Length", C.c_double), ("BendingAngle", C.c_double), ("EntranceAngle", C.c_double),
        ("ExitAngle", C.c_double), ("FringeInt1", C.c_double), ("FringeInt2", C.c_double),
        ("FullGap", C.c_double), ("FieldScaling", C.c_double), ("Voltage", C.c_double),
        ("Frequency", C.c_double), ("TimeLag", C.c_double), ("PhaseLag", C.c_double),
        ("bax", C.c_double), ("bay", C.c_double), ("Energy", C.c_double),
        ("PolynomA", _dp), ("PolynomB", _dp), ("fringeIntM0", _dp), ("fringeIntP0", _dp),
        ("R1", _dp), ("R2", _dp), ("T1", _dp), ("T2", _dp), ("RApertures", _dp),
        ("EApertures", _dp), ("KickAngle", _dp), ("Limits", _dp),
    ]


_ARRAYS = ("PolynomA", "PolynomB", "fringeIntM0", "fringeIntP0", "R1", "R2", "T1", "T2",
           "RApertures", "EApertures", "KickAngle", "Limits")


class _Lib:
    def __init__(self, path: str, prefix: str):
        self.lib = C.CDLL(path)
        self.atpass = getattr(self.lib, prefix + "atpass")
        self.atpass.restype = C.c_int
        self.atpass.argtypes = [C.POINTER(OracleElem), C.c_int32, _dp, C.c_int32, C.c_int32,
                                C.POINTER(C.c_uint32), C.c_int32, C.c_int32, C.c_double,
                                C.c_double, C.c_int32, C.c_int32, C.c_int32, _dp,
                                C.POINTER(C.c_uint8), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint32), _dp, C.c_int32]
        self.elempass = getattr(self.lib, prefix + "elempass")
        self.elempass.restype = C.c_int
        self.elempass.argtypes = [C.POINTER(OracleElem), _dp, C.c_int32, C.c_double, C.c_double]
        self.err = getattr(self.lib, prefix + "last_error")
        self.err.restype = C.c_char_p


_libs: dict = {}


def available(impl: str = "oracle") -> bool:
    name = "liboracle.so" if impl == "oracle" else os.path.join("_ref", "libref_driver.so")
    return os.path.exists(os.path.join(HERE, name))


def _lib(impl: str) -> _Lib:
    if impl not in _libs:
        if impl == "oracle":
            _libs[impl] = _Lib(os.path.join(HERE, "liboracle.so"), "oracle_")
        elif impl == "ref":
            lib = _Lib(os.path.join(REF_DIR, "libref_driver.so"), "ref_")
            setdir = lib.lib.ref_set_library_dir
            setdir.argtypes = [C.c_char_p]
            setdir(REF_DIR.encode())
            _libs[impl] = lib
        else:
            raise ValueError(impl)
    return _libs[impl]


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def make_elem(e, keep: list) -> OracleElem:
    """Flatten one element object into the oracle record, applying the
    reference defaults of each trackFunction (atintegrators/<Pass>.c)."""
    pm = e.PassMethod
    if pm not in PASS_IDS:
        raise RuntimeError(f"PassMethod {pm}: library, module or trackFunction not found")
    s = OracleElem()
    s.pass_ = PASS_IDS[pm]
    g = lambda n, d=0.0: getattr(e, n, d)  # noqa: E731
    s.Length = float(g("Length", 0.0))
    s.MaxOrder = int(g("MaxOrder", 0))
    s.NumIntSteps = int(g("NumIntSteps", 10))
    ba = g("BendingAngle", 0.0)
    if pm == "ThinMPolePass":
        arr = np.atleast_1d(np.asarray(ba, dtype=np.float64)) if hasattr(e, "BendingAngle") else []
        s.bax = float(arr[0]) if len(arr) >= 1 else 0.0
        s.bay = float(arr[1]) if len(arr) >= 2 else 0.0
    else:
        s.BendingAngle = float(ba)
    s.EntranceAngle = float(g("EntranceAngle"))
    s.ExitAngle = float(g("ExitAngle"))
    s.FringeBendEntrance = int(g("FringeBendEntrance", 1))
    s.FringeBendExit = int(g("FringeBendExit", 1))
    s.FringeInt1 = float(g("FringeInt1"))
    s.FringeInt2 = float(g("FringeInt2"))
    s.FullGap = float(g("FullGap"))
    s.FringeQuadEntrance = int(g("FringeQuadEntrance", 0))
    s.FringeQuadExit = int(g("FringeQuadExit", 0))
    s.FieldScaling = float(g("FieldScaling", 1.0))
    s.Voltage = float(g("Voltage"))
    s.Frequency = float(g("Frequency"))
    s.TimeLag = float(g("TimeLag"))
    s.PhaseLag = float(g("PhaseLag"))
    s.Energy = float(g("Energy", 0.0))
    for name in _ARRAYS:
        v = getattr(e, name, None)
        if v is None:
            continue
        a = np.asarray(v, dtype=np.float64)
        if name in ("PolynomA", "PolynomB"):
            a = a.ravel()
            need = max(int(g("MaxOrder", 0)) + 1, 2)
            if a.size < need:
                a = np.pad(a, (0, need - a.size))
        elif name in ("R1", "R2"):
            a = a.reshape((6, 6))
        a = np.require(a, np.float64, ["F", "A"]).ravel(order="F").copy()
        keep.append(a)
        setattr(s, name, _ptr(a))
    return s


def make_line(lattice) -> tuple:
    keep: list = []
    n = len(lattice)
    arr = (OracleElem * max(n, 1))()
    for i, e in enumerate(lattice):
        arr[i] = make_elem(e, keep)
    return arr, keep


def _energy_particle(lattice, energy, particle):
    if energy is None:
        energy = getattr(lattice, "energy", 0.0)
    if particle is None:
        particle = getattr(lattice, "particle", None)
    rest = float(getattr(particle, "rest_energy", 0.0)) if particle is not None else 0.0
    return float(energy or 0.0), rest


def atpass(lattice, rin, nturns, refpts=None, turn=0, energy=None, particle=None,
           losses=False, start_elem=0, end_elem=0, impl="oracle", nthreads=0):
    """Track in place; returns rout (6,N,R,T) or (rout, lossdict) like at.c."""
    if rin.shape[0] != 6 or rin.dtype != np.float64 or not rin.flags.f_contiguous:
        raise ValueError("rin must be a 6xN float64 Fortran-ordered array")
    lib = _lib(impl)
    npart = rin.size // 6
    refs = np.zeros(0, np.uint32) if refpts is None else np.ascontiguousarray(refpts, np.uint32)
    line, keep = make_line(lattice)
    e, rest = _energy_particle(lattice, energy, particle)
    rout = np.full((6, npart, refs.size, nturns), np.nan, order="F")
    lf = np.zeros(npart, np.uint8)
    lt = np.zeros(npart, np.uint32)
    le = np.zeros(npart, np.uint32)
    lc = np.zeros((6, npart), order="F")
    rc = lib.atpass(line, len(lattice), _ptr(rin), npart, nturns,
                    refs.ctypes.data_as(C.POINTER(C.c_uint32)), refs.size, turn, e, rest,
                    int(losses), start_elem, end_elem, _ptr(rout),
                    lf.ctypes.data_as(C.POINTER(C.c_uint8)),
                    lt.ctypes.data_as(C.POINTER(C.c_uint32)),
                    le.ctypes.data_as(C.POINTER(C.c_uint32)), _ptr(lc), nthreads)
    if rc:
        msg = lib.err().decode()
        raise (ValueError if rc in (1, 3) else RuntimeError)(msg)
    del keep
    if losses:
        return rout, {"islost": lf.astype(bool), "turn": lt, "elem": le, "coord": lc}
    return rout


def elempass(element, rin, energy=0.0, particle=None, impl="oracle"):
    if rin.shape[0] != 6 or rin.dtype != np.float64 or not rin.flags.f_contiguous:
        raise ValueError("rin must be a 6xN float64 Fortran-ordered array")
    lib = _lib(impl)
    keep: list = []
    s = make_elem(element, keep)
    rest = float(getattr(particle, "rest_energy", 0.0)) if particle is not None else 0.0
    rc = lib.elempass(C.byref(s), _ptr(rin), rin.size // 6, float(energy or 0.0), rest)
    if rc:
        raise (ValueError if rc in (1, 3) else RuntimeError)(lib.err().decode())
    return rin


def tangent_ref(element, r6) -> np.ndarray:
    """6x6 Jacobian (row-major) of one Str/BndMPole element at r6, read off
    the reference's own diffusion-matrix propagation (diff_drift.c,
    diff_str_kick.c, diff_bnd_kick.c, diff_bend_fringe.c) through
    ref_tangent in ref_driver.c. Elements must have no T1/T2, R1/R2 or
    quadrupole fringe, which the cores do not propagate."""
    lib = _lib("ref")
    fn = lib.lib.ref_tangent
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(OracleElem), _dp, _dp]
    keep: list = []
    s = make_elem(element, keep)
    r = np.ascontiguousarray(r6, dtype=np.float64).reshape(6)
    m = np.zeros((6, 6), np.float64)
    rc = fn(C.byref(s), _ptr(r), _ptr(m))
    if rc:
        raise RuntimeError(lib.err().decode())
    del keep
    return m


def diffusion_ref(element, orbit, energy, particle=None) -> np.ndarray:
    """Ohmi diffusion matrix (6x6) of one element from its entrance orbit,
    computed by the reference cores as at_diffmatrix does (ref_diffusion
    in ref_driver.c)."""
    lib = _lib("ref")
    fn = lib.lib.ref_diffusion
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(OracleElem), _dp, _dp, C.c_double, C.c_double]
    keep: list = []
    s = make_elem(element, keep)
    o = np.ascontiguousarray(orbit, dtype=np.float64).reshape(6)
    b = np.zeros(36, np.float64)
    rest = float(getattr(particle, "rest_energy", 0.0)) if particle is not None else 0.0
    rc = fn(C.byref(s), _ptr(o), _ptr(b), float(energy), rest)
    if rc:
        raise RuntimeError(lib.err().decode())
    del keep
    return b.reshape(6, 6).T.copy()
