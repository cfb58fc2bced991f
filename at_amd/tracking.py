"""Tracking API with pyat's call surface (pyat/at/tracking/track.py).

``lattice_track``, ``element_track``, ``internal_lpass``/``internal_epass``,
``MPMode``, ``gpu_info`` and ``gpu_core_count`` keep the reference names,
keyword arguments, return values and errors. Every mode runs the HIP engine
(libatx.so): the reference's ``use_mp=MPMode.GPU`` switch
(track.py:115-127) is this package's only compute path, and ``gpu_pool``
selects the device. There is no CPU fallback.
"""

from __future__ import annotations

import functools
from collections.abc import Iterable
from enum import Enum

import numpy as np

from . import engine
from .lattice import AtError, Element, Lattice, Particle

__all__ = [
    "MPMode", "lattice_track", "element_track", "internal_lpass", "internal_epass",
    "lattice_pass", "element_pass", "gpu_info", "gpu_core_count", "gpupass", "atpass",
    "get_uint32_index", "End", "diffusion_matrix", "diffusion_matrices", "element_matrices",
    "patpass",
    "internal_plpass",
]


class MPMode(Enum):
    """Multi-processing mode (track.py:32-38)."""

    CPU = 1
    GPU = 2


class _End:
    def __repr__(self):
        return "End"


End = _End()

DIMENSION_ERROR = "Input to lattice_pass() must be a 6xN array."


def get_uint32_index(lattice, refpts) -> np.ndarray:
    """Normalise refpts to a sorted uint32 index array
    (pyat/at/lattice/utils.py:417): None -> [], End -> [len], int, bool
    masks, and sequences; negative indices count from len+1."""
    n = len(lattice)
    if refpts is None:
        return np.zeros(0, np.uint32)
    if refpts is End:
        return np.array([n], np.uint32)
    arr = np.asarray(refpts)
    if arr.dtype == bool:
        idx = np.flatnonzero(arr)
    else:
        idx = np.atleast_1d(arr).astype(np.int64)
        idx = np.where(idx < 0, idx + n + 1, idx)
    if np.any(idx < 0) or np.any(idx > n):
        raise IndexError(f"refpts out of range [0..{n}]")
    return np.unique(idx).astype(np.uint32)


def fortran_align(func):
    """Ensure r_in is Fortran-ordered, copying back into r_in afterwards
    (pyat/at/tracking/utils.py:90-115)."""

    @functools.wraps(func)
    def wrapper(lattice, r_in, *args, **kwargs):
        assert r_in.shape[0] == 6 and r_in.ndim in (1, 2), DIMENSION_ERROR
        if r_in.flags.f_contiguous:
            return func(lattice, r_in, *args, **kwargs)
        r_fin = np.asfortranarray(r_in)
        r_out = func(lattice, r_fin, *args, **kwargs)
        r_in[:] = r_fin[:]
        return r_out

    return wrapper


class _Cache:
    """Lattice reuse across calls (at.c:445-512 keep_lattice/reuse, and the
    static Lattice of atgpu/PyInterface.cpp:135-137)."""

    lattice_id = None
    device_lattice: engine.DeviceLattice | None = None
    last_turn = 0


def _energy_particle(lattice, kwargs):
    energy = kwargs.pop("energy", None)
    particle = kwargs.pop("particle", None)
    if energy is None:
        energy = getattr(lattice, "energy", 0.0)
    if particle is None:
        particle = getattr(lattice, "particle", None)
    if isinstance(particle, str):
        particle = Particle(particle)
    rest = float(getattr(particle, "rest_energy", 0.0)) if particle is not None else 0.0
    charge = float(getattr(particle, "charge", -1.0)) if particle is not None else -1.0
    return float(energy or 0.0), rest, charge


def gpupass(lattice, r_in, nturns: int, refpts=None, turn: int | None = None,
            energy=None, particle=None, keep_counter: bool = False, reuse: bool = False,
            losses: bool = False, bunch_spos=None, bunch_currents=None,
            gpu_pool: list[int] | None = None, integrator: int = 4, verbose: bool = False,
            start_elem: int | None = None, end_elem: int | None = None, strict: bool = False,
            repack: bool = False, **_ignored):
    """Engine entry with the keyword set of at.tracking.gpu.gpupass
    (pyat/at/tracking/gpu.pyi:9-28; atgpu/PyInterface.cpp:139-197).
    ``repack`` (calls without refpts): turn blocks of 1, 1, 2, 4, ... turns
    with the survivors packed into the leading workgroups between them
    (ATX_REPACK), identical results. It pays where most particles are lost
    early (dynamic-aperture grids: the acceptance drivers set it); on a
    beam that loses a few percent over the run the block boundaries cost
    more than the packing saves (1.6 % on the large ring's 100-turn
    calls)."""
    if not isinstance(lattice, list):
        raise TypeError("lattice must be a list of elements")
    if not isinstance(r_in, np.ndarray):
        raise TypeError("r_in must be a numpy array")
    if not isinstance(nturns, (int, np.integer)):
        raise TypeError("nturns must be an integer")
    if gpu_pool is not None:
        if len(gpu_pool) == 0:
            raise RuntimeError("at_gpupass() gpu_pool is empty")
        if len(gpu_pool) > 1:
            raise RuntimeError("at_gpupass() Multi GPU support not implemented")
        engine.set_device(int(gpu_pool[0]))
    kw = {}
    if energy is not None:
        kw["energy"] = energy
    if particle is not None:
        kw["particle"] = particle
    e, rest, charge = _energy_particle(lattice, kw)
    if not (reuse and _Cache.lattice_id == id(lattice) and _Cache.device_lattice is not None):
        _Cache.device_lattice = None
        _Cache.lattice_id = None
        _Cache.device_lattice = engine.DeviceLattice(lattice)
    dl = _Cache.device_lattice
    if getattr(dl, "integrator", 4) != integrator:
        if verbose and reuse:
            print("Warning, lattice is recreated when integrator type is changed")
        dl.set_integrator(integrator)
    if verbose:
        print(f"Lattice created [{dl.nelem} elements]; tracking {r_in.size // 6} particles "
              f"for {nturns} turns")
    refs = np.zeros(0, np.uint32) if refpts is None else np.asarray(refpts, np.uint32)
    t0 = _Cache.last_turn if keep_counter else (0 if turn is None else int(turn))
    res = dl.track(r_in, int(nturns), refs, turn=t0, energy=e, rest_energy=rest, charge=charge,
                   losses=losses, start_elem=int(start_elem or 0), end_elem=int(end_elem or 0),
                   strict=strict, repack=bool(repack))
    _Cache.lattice_id = id(lattice)
    _Cache.last_turn = t0 + int(nturns)
    return res


# The engine's atpass has the signature of pyat's at.tracking.atpass.atpass
# (pyat/at.c:340-345), for code that calls the extension directly.
def atpass(line, rin, nturns, refpts=None, turn=0, energy=None, particle=None,
           keep_counter=False, reuse=False, omp_num_threads=0, losses=False,
           bunch_spos=None, bunch_currents=None, start_elem=0, end_elem=0,
           harmonic_number=None, fillpattern=None, strict=False):
    if not isinstance(line, list):
        raise TypeError("line must be a list")
    if not isinstance(rin, np.ndarray):
        raise TypeError("rin must be a numpy array")
    if not isinstance(nturns, (int, np.integer)):
        raise TypeError("nturns must be an integer")
    if rin.shape[0] != 6:
        raise ValueError("rin is not 6D")
    if rin.dtype != np.float64:
        raise ValueError("rin is not a double array")
    if not rin.flags.f_contiguous:
        raise ValueError("rin is not Fortran-aligned")
    if refpts is not None and np.asarray(refpts).dtype != np.uint32:
        raise ValueError("refpts is not a uint32 array")
    return gpupass(line, rin, nturns, refpts=refpts, turn=turn, energy=energy,
                   particle=particle, keep_counter=keep_counter, reuse=reuse, losses=losses,
                   start_elem=start_elem, end_elem=end_elem, strict=strict)


@fortran_align
def _element_pass(element: Element, r_in, energy=None, particle=None, strict=False, **kwargs):
    kw = {}
    if energy is not None:
        kw["energy"] = energy
    if particle is not None:
        kw["particle"] = particle
    e, rest, _ = _energy_particle(None, kw)
    dl = engine.DeviceLattice([element])
    try:
        return dl.elempass(0, r_in, energy=e, rest_energy=rest, strict=strict)
    finally:
        dl.close()


@fortran_align
def _lattice_pass(lattice: list, r_in, nturns: int = 1, refpts=End, no_varelem: bool = True,
                  seed: int | None = None, use_gpu: bool = True, **kwargs):
    """internal_lpass (track.py:96-129)."""
    kwargs["reuse"] = kwargs.pop("keep_lattice", False)
    refs = get_uint32_index(lattice, refpts)
    kwargs.pop("harmonic_number", None)
    kwargs.pop("fillpattern", None)
    kwargs.pop("omp_num_threads", None)
    kwargs.pop("use_mp", None)
    kwargs.pop("pool_size", None)
    kwargs.pop("start_method", None)
    return gpupass(lattice, r_in, nturns, refpts=refs, **kwargs)


@fortran_align
def _plattice_pass(lattice: list, r_in, nturns: int = 1, refpts=End, seed: int | None = None,
                   pool_size: int | None = None, start_method: str | None = None, **kwargs):
    """internal_plpass (track.py:133-180). The reference splits the particles
    over a process pool; one device launch tracks them all here, so
    pool_size and start_method are accepted and unused."""
    return _lattice_pass(lattice, r_in, nturns, refpts, seed=seed, **kwargs)


def patpass(lattice, r_in, nturns: int = 1, refpts=End, pool_size: int | None = None,
            start_method: str | None = None, **kwargs):
    """Parallel lattice_pass (tracking/deprecated.py:190-226): r_in is
    tracked in place and rout is returned."""
    return _plattice_pass(lattice, r_in, nturns=nturns, refpts=refpts, pool_size=pool_size,
                          start_method=start_method, **kwargs)


_mp_table = {
    False: (False, False),
    True: (True, False),
    MPMode.CPU: (True, False),
    MPMode.GPU: (False, True),
}


def lattice_track(lattice: Iterable[Element], r_in, nturns: int = 1, refpts=End,
                  in_place: bool = False, **kwargs):
    """Track particles through a lattice (track.py:181-382).

    Returns ``(r_out, trackparam, trackdata)`` with r_out of shape
    (6, N, R, T) and ``trackdata["loss_map"]`` a recarray
    (islost, turn, elem, coord)."""
    use_mp = kwargs.pop("use_mp", False)
    if use_mp not in _mp_table:
        raise AtError(f"Invalid use_mp value: {use_mp!r}")
    trackdata: dict = {}
    trackparam: dict = {}
    try:
        npart = np.shape(r_in)[1]
    except IndexError:
        npart = 1
    for kw in ("energy", "particle"):
        if kw in kwargs:
            trackparam[kw] = kwargs.get(kw)
    trackparam["npart"] = npart
    if not in_place:
        r_in = r_in.copy()
    if not isinstance(lattice, list):
        lattice = list(lattice)
    kwargs.pop("unfold_beam", None)
    ldtype = [("islost", np.bool_), ("turn", np.uint32), ("elem", np.uint32),
              ("coord", np.float64, (6,))]
    loss_map = np.recarray((npart,), ldtype)
    if "turn" in kwargs:
        trackparam["turn"] = kwargs.get("turn")
    trackparam.update({"refpts": get_uint32_index(lattice, refpts), "nturns": nturns})
    kwargs.pop("start_method", None)
    kwargs.pop("pool_size", None)
    rout = _lattice_pass(lattice, r_in, nturns=nturns, refpts=refpts, no_varelem=False, **kwargs)
    if kwargs.get("losses", False):
        rout, lm = rout
        lm["coord"] = lm["coord"].T
        for k, v in lm.items():
            loss_map[k] = v
    trackdata["loss_map"] = loss_map
    trackparam["rout"] = r_in
    return rout, trackparam, trackdata


def element_track(element: Element, r_in, in_place: bool = False, **kwargs):
    """Track particles through one element (track.py:385-420)."""
    if not in_place:
        r_in = r_in.copy()
    return _element_pass(element, r_in, **kwargs)


def lattice_pass(lattice, r_in, nturns: int = 1, refpts=End, **kwargs):
    """Legacy form returning r_out only, tracking r_in in place."""
    return lattice_track(lattice, r_in, nturns, refpts, in_place=True, **kwargs)[0]


def element_pass(element, r_in, **kwargs):
    return element_track(element, r_in, in_place=True, **kwargs)


def _diffusion_energy(energy, particle):
    # at_diffmatrix takes the energy and particle only from its keywords
    # (pyat/at.c:755-800): without a particle the rest energy stays 0 and
    # the radiating cores take gamma as E / m_e (atelem.c:363-370)
    if isinstance(particle, str):
        particle = Particle(particle)
    rest = float(getattr(particle, "rest_energy", 0.0)) if particle is not None else 0.0
    return float(energy or 0.0), rest


def diffusion_matrix(element: Element, rin, *, energy=None, particle=None) -> np.ndarray:
    """Ohmi diffusion matrix of one element from its entrance orbit, with
    the argument checks of at_diffmatrix (pyat/at.c:755-831): rin is a
    Fortran-ordered (6, 1) double array. Non-radiating passes give zeros."""
    if not isinstance(rin, np.ndarray):
        raise TypeError("rin must be a numpy array")
    if rin.shape[0] != 6:
        raise ValueError("rin is not 6D")
    if rin.dtype != np.float64:
        raise ValueError("rin is not a double array")
    if not rin.flags.f_contiguous:
        raise ValueError("rin is not Fortran-aligned")
    if rin.size != 6:
        raise ValueError("Number of particles should be 1")
    e, rest = _diffusion_energy(energy, particle)
    dl = engine.DeviceLattice([element])
    try:
        return dl.diffusion_matrices(rin.reshape(6, 1), e, rest)[0]
    finally:
        dl.close()


def diffusion_matrices(lattice, orbits, *, energy=None, particle=None,
                       keep_lattice: bool = False) -> np.ndarray:
    """Diffusion matrices of every element of a lattice in one launch:
    orbits (6, nelem) holds each element's entrance orbit; returns
    (nelem, 6, 6). Element by element this equals diffusion_matrix, which
    pyat's _dmatr calls in a loop (physics/radiation.py:98-115). The device
    lattice is shared with gpupass and reused with keep_lattice."""
    if not isinstance(lattice, list):
        lattice = list(lattice)
    orbits = np.asfortranarray(orbits, dtype=np.float64)
    e, rest = _diffusion_energy(energy, particle)
    if not (keep_lattice and _Cache.lattice_id == id(lattice)
            and _Cache.device_lattice is not None):
        _Cache.device_lattice = None
        _Cache.lattice_id = None
        _Cache.device_lattice = engine.DeviceLattice(lattice)
        _Cache.lattice_id = id(lattice)
    return _Cache.device_lattice.diffusion_matrices(orbits, e, rest)


def element_matrices(lattice, orbits=None, *, step: float, energy=None, particle=None,
                     strict: bool = False, keep_lattice: bool = False) -> np.ndarray:
    """6x6 matrix of every element of a lattice around its entrance orbit
    (orbits (6, nelem), default 0) by central differences of ``step``, in one
    launch; returns (nelem, 6, 6). Element by element this equals
    find_elem_m66 (pyat/at/physics/matrix.py:171-202), which pyat's _dmatr
    calls in a loop (physics/radiation.py:78-90). Energy and particle default
    to the lattice's. The device lattice is shared with gpupass and reused
    with keep_lattice."""
    kw = {}
    if energy is not None:
        kw["energy"] = energy
    if particle is not None:
        kw["particle"] = particle
    e, rest, _ = _energy_particle(lattice, kw)
    elems = list(lattice)
    if orbits is None:
        orbits = np.zeros((6, len(elems)))
    orbits = np.asfortranarray(orbits, dtype=np.float64)
    if not (keep_lattice and _Cache.lattice_id == id(lattice)
            and _Cache.device_lattice is not None):
        _Cache.device_lattice = None
        _Cache.lattice_id = None
        _Cache.device_lattice = engine.DeviceLattice(elems)
        _Cache.lattice_id = id(lattice)
    return _Cache.device_lattice.element_matrices(orbits, step, e, rest, strict=strict)


def gpu_info() -> list:
    """[name, architecture, compute units, platform] per device
    (track.py:444-461)."""
    try:
        n = engine.device_count()
    except ImportError:
        return []
    return [engine.device_info(i) for i in range(n)]


def gpu_core_count(gpuPool: list[int] | None) -> int:
    """Number of compute units in the selected devices (track.py:423-441)."""
    gpus = gpu_info()
    nbgpu = len(gpus)
    if gpuPool is None:
        gpuPool = [0]
    nbcore = 0
    for g in gpuPool:
        if g < 0 or g >= nbgpu:
            raise AtError(f"Invalid GPU id {g}, must be in [0..{nbgpu - 1}]")
        nbcore += gpus[g][2]
    return nbcore


internal_lpass = _lattice_pass
internal_epass = _element_pass
internal_plpass = _plattice_pass
Lattice.track = lattice_track
Element.track = element_track
