"""ctypes binding of libatx.so (C ABI in include/atx.h).

``DeviceLattice`` turns element objects into an engine lattice handle by
setting every numeric attribute by name, exactly as atgpu's PyInterface
feeds PassMethodFactory (atgpu/PyInterface.cpp:404-480). The library is the
only compute path: importing this module raises when it is missing.
"""

from __future__ import annotations

import atexit
import ctypes as C
import os
import threading

import numpy as np

from .lattice import AtError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ATX_LIBRARY", os.path.join(_HERE, "libatx.so"))

ATX_OK = 0
ATX_E_VALUE = 1
ATX_E_ATTRIBUTE = 2
ATX_E_RUNTIME = 3
ATX_E_NODEVICE = 4

ATX_DEVICE_POINTERS = 0x1
ATX_STRICT = 0x2
ATX_KEEP_COUNTER = 0x4
ATX_SOA = 0x8
ATX_CONTINUE_LOSSES = 0x10
ATX_REPACK = 0x20

# Exported symbols, as declared in include/atx.h.
EXPORTS = (
    "atx_last_error", "atx_lattice_create", "atx_lattice_destroy", "atx_lattice_add_element",
    "atx_element_set_double", "atx_element_set_int", "atx_element_set_array",
    "atx_lattice_num_elements", "atx_lattice_check", "atx_track", "atx_elempass",
    "atx_track_batch", "atx_lattice_turn_counter", "atx_device_count", "atx_device_info",
    "atx_set_device", "atx_lattice_set_integrator", "atx_adjoint", "atx_element_set_array_at",
    "atx_diffusion_matrices", "atx_element_matrices", "atx_adjoint_batch", "atx_shutdown", "atx_get_device",
    "atx_device_alloc", "atx_device_free", "atx_host_alloc", "atx_host_free", "atx_memcpy",
    "atx_stream_create", "atx_stream_destroy", "atx_stream_synchronize", "atx_stream_wait_event",
    "atx_event_create", "atx_event_destroy", "atx_event_record", "atx_event_query",
    "atx_event_synchronize", "atx_event_elapsed", "atx_lattice_set_launch_shape",
    "atx_track_batch_noalloc",
)


class BatchItem(C.Structure):
    """struct atx_batch_item of include/atx.h."""

    _fields_ = [("lattice", C.c_void_p), ("r", C.c_void_p), ("num_particles", C.c_uint32),
                ("pad", C.c_uint32), ("lost_flag", C.c_void_p), ("lost_turn", C.c_void_p),
                ("lost_elem", C.c_void_p), ("lost_coord", C.c_void_p), ("nturns", C.c_int32),
                ("turn", C.c_int32)]

class LaunchShape(C.Structure):
    """struct atx_launch_shape of include/atx.h (zero fields: engine default)."""

    _fields_ = [("block", C.c_int32), ("ppt", C.c_int32), ("unroll", C.c_int32),
                ("pad", C.c_int32)]


def launch_shape(block: int = 0, ppt: int = 0, unroll: int = 0):
    """A LaunchShape, or None (the engine's default) when every field is 0."""
    if not (block or ppt or unroll):
        return None
    return LaunchShape(int(block), int(ppt), int(unroll), 0)


class AdjointItem(C.Structure):
    """struct atx_adjoint_item of include/atx.h."""

    _fields_ = [("states", C.c_void_p), ("lam", C.c_void_p), ("nturns", C.c_int32),
                ("turn", C.c_int32), ("num_particles", C.c_uint32), ("accumulate", C.c_int32)]


_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_u8p = C.POINTER(C.c_uint8)

_lib = None
_lib_lock = threading.Lock()


def _bind_runtime() -> None:
    """Keep one HIP runtime in the process, without importing torch.

    PyTorch-ROCm ships its own libamdhip64 (soname libamdhip64.so.7, which
    torch loads from its lib directory as "libamdhip64.so"). A process in
    which both the engine and torch run must hold one runtime: torch's
    device pointers and streams are then valid in the engine, and torch,
    imported after the engine, does not start a second runtime (which fails
    to initialise: "No HIP GPUs are available"). So, when torch is
    installed, that library file is loaded (globally, by path, before
    libatx.so resolves its libamdhip64.so.7) -- the runtime only, not torch.
    ``ATX_HIP_RUNTIME=system`` skips this and uses the system ROCm runtime
    (for processes that never use torch); without torch installed that is
    the runtime anyway."""
    import importlib.util
    import sys

    if "torch" in sys.modules or os.environ.get("ATX_HIP_RUNTIME", "") == "system":
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for base in spec.submodule_search_locations:
        lib = os.path.join(base, "lib", "libamdhip64.so")
        if os.path.exists(lib):
            C.CDLL(lib, mode=C.RTLD_GLOBAL)
            return


def load_library(path: str | None = None) -> C.CDLL:
    """Load libatx.so and declare its signatures. Raises when it is absent:
    there is no fallback compute path."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise ImportError(f"at_amd engine library not built: {p} (run python -m at_amd.build)")
        _bind_runtime()
        lib = C.CDLL(p)
        lib.atx_last_error.restype = C.c_char_p
        lib.atx_lattice_create.argtypes = [C.POINTER(C.c_void_p)]
        lib.atx_lattice_destroy.argtypes = [C.c_void_p]
        lib.atx_lattice_add_element.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        lib.atx_element_set_double.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        lib.atx_element_set_int.argtypes = [C.c_void_p, C.c_char_p, C.c_int64]
        lib.atx_element_set_array.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_int64]
        lib.atx_lattice_num_elements.argtypes = [C.c_void_p]
        lib.atx_lattice_check.argtypes = [C.c_void_p, C.c_double, C.c_double]
        lib.atx_track.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int32, _u32p, C.c_uint32,
                                  C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32,
                                  C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]
        lib.atx_elempass.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint32, C.c_double,
                                     C.c_double, C.c_void_p, C.c_uint32]
        lib.atx_track_batch.argtypes = [C.POINTER(BatchItem), C.c_int32, C.c_int32, C.c_int32,
                                        C.c_double, C.c_double, C.c_int32,
                                        C.POINTER(LaunchShape), C.c_void_p, C.c_uint32]
        lib.atx_lattice_set_launch_shape.argtypes = [C.c_void_p, C.POINTER(LaunchShape)]
        lib.atx_track_batch_noalloc.argtypes = lib.atx_track_batch.argtypes
        lib.atx_lattice_turn_counter.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        lib.atx_lattice_set_integrator.argtypes = [C.c_void_p, C.c_int32]
        lib.atx_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_uint32,
                                    C.c_double, C.c_double, C.c_void_p,
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_void_p,
                                    C.c_int32, C.c_void_p]
        lib.atx_adjoint_batch.argtypes = [C.c_void_p, C.POINTER(AdjointItem), C.c_int32,
                                          C.c_double, C.c_double, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_void_p, C.c_int32,
                                          C.c_void_p]
        lib.atx_element_set_array_at.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, _dp,
                                                 C.c_int64]
        lib.atx_element_matrices.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                             C.c_double, C.c_double, C.c_void_p, C.c_uint32]
        lib.atx_diffusion_matrices.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                               C.c_double, C.c_void_p, C.c_uint32]
        lib.atx_device_count.argtypes = [C.POINTER(C.c_int)]
        lib.atx_device_info.argtypes = [C.c_int, C.c_char_p, C.c_int, C.c_char_p, C.c_int,
                                        C.POINTER(C.c_int), C.c_char_p, C.c_int]
        lib.atx_set_device.argtypes = [C.c_int]
        lib.atx_shutdown.argtypes = []
        vp, pvp = C.c_void_p, C.POINTER(C.c_void_p)
        lib.atx_get_device.argtypes = [C.POINTER(C.c_int)]
        lib.atx_device_alloc.argtypes = [pvp, C.c_uint64]
        lib.atx_device_free.argtypes = [vp]
        lib.atx_host_alloc.argtypes = [pvp, C.c_uint64]
        lib.atx_host_free.argtypes = [vp]
        lib.atx_memcpy.argtypes = [vp, vp, C.c_uint64, vp]
        lib.atx_stream_create.argtypes = [pvp]
        lib.atx_stream_destroy.argtypes = [vp]
        lib.atx_stream_synchronize.argtypes = [vp]
        lib.atx_stream_wait_event.argtypes = [vp, vp]
        lib.atx_event_create.argtypes = [pvp]
        lib.atx_event_destroy.argtypes = [vp]
        lib.atx_event_record.argtypes = [vp, vp]
        lib.atx_event_query.argtypes = [vp, C.POINTER(C.c_int)]
        lib.atx_event_synchronize.argtypes = [vp]
        lib.atx_event_elapsed.argtypes = [C.POINTER(C.c_float), vp, vp]
        for name in EXPORTS:
            getattr(lib, name).restype = C.c_char_p if name == "atx_last_error" else C.c_int
        _lib = lib
        # release the engine's graphs, events and pinned buffers while the
        # HIP runtime is still whole: Python's exit handlers run before the
        # shared libraries' destructors
        atexit.register(lib.atx_shutdown)
        return lib


def _raise(rc: int, lib) -> None:
    msg = lib.atx_last_error().decode(errors="replace")
    if rc == ATX_E_VALUE:
        raise ValueError(msg)
    if rc == ATX_E_ATTRIBUTE:
        raise AttributeError(msg)
    raise RuntimeError(msg)


def _check(rc: int, lib) -> None:
    if rc != ATX_OK:
        _raise(rc, lib)


# Attributes that are never sent to the engine (names and bookkeeping).
_SKIP = {"FamName", "PassMethod", "Class"}


def _send_attribute(lib, h, name: str, value) -> None:
    if name in _SKIP or name.startswith("_"):
        return
    bname = name.encode()
    if isinstance(value, (bool, np.bool_)):
        _check(lib.atx_element_set_int(h, bname, int(value)), lib)
    elif isinstance(value, (int, np.integer)):
        _check(lib.atx_element_set_int(h, bname, int(value)), lib)
    elif isinstance(value, (float, np.floating)):
        _check(lib.atx_element_set_double(h, bname, float(value)), lib)
    elif isinstance(value, np.ndarray) and value.dtype.kind in "fiub":
        a = np.asarray(value, dtype=np.float64)
        a = np.ascontiguousarray(a.ravel(order="F"))
        _check(lib.atx_element_set_array(h, bname, a.ctypes.data_as(_dp), a.size), lib)
    elif isinstance(value, (list, tuple)) and all(isinstance(v, (int, float)) for v in value):
        a = np.ascontiguousarray(np.asarray(value, dtype=np.float64))
        _check(lib.atx_element_set_array(h, bname, a.ctypes.data_as(_dp), a.size), lib)
    # other attribute types (strings, objects) carry no pass-method parameter


class DeviceLattice:
    """Engine lattice handle built from a sequence of elements."""

    def __init__(self, elements):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.atx_lattice_create(C.byref(h)), self.lib)
        self.h = h
        self.nelem = 0
        self.integrator = 4
        try:
            for i, e in enumerate(elements):
                pm = getattr(e, "PassMethod", None)
                if pm is None:
                    raise AttributeError(f"Element #{i} has no PassMethod")
                fam = str(getattr(e, "FamName", ""))
                rc = self.lib.atx_lattice_add_element(h, pm.encode(), fam.encode())
                if rc != ATX_OK:
                    _raise(rc, self.lib)
                for name, value in vars(e).items():
                    _send_attribute(self.lib, h, name, value)
                self.nelem += 1
        except Exception:
            self.close()
            raise

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.atx_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_launch_shape(self, block: int = 0, ppt: int = 0, unroll: int = 0) -> None:
        """Launch shape of this lattice's atx_track/atx_elempass calls (all 0:
        the engine's default)."""
        sh = launch_shape(block, ppt, unroll)
        _check(self.lib.atx_lattice_set_launch_shape(self.h, None if sh is None else C.byref(sh)),
               self.lib)

    def set_integrator(self, integrator: int) -> None:
        """Integrator type 1-6 of atgpu's SymplecticIntegrator (4 = CPU)."""
        _check(self.lib.atx_lattice_set_integrator(self.h, int(integrator)), self.lib)
        self.integrator = int(integrator)

    def set_array_at(self, index: int, name: str, values) -> None:
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel(order="F"))
        _check(self.lib.atx_element_set_array_at(self.h, int(index), name.encode(),
                                                 a.ctypes.data_as(_dp), a.size), self.lib)

    def adjoint(self, states_ptr: int, nturns: int, turn: int, npart: int, energy: float,
                rest_energy: float, lam_ptr: int, slots: np.ndarray, coeffs: np.ndarray,
                grad_ptr: int, nparams: int, stream: int | None = None) -> None:
        """Reverse-mode pass over one block of turns (atx_adjoint)."""
        slots = np.ascontiguousarray(slots, np.int32)
        coeffs = np.ascontiguousarray(coeffs, np.int32)
        if slots.size != self.nelem or coeffs.size != self.nelem:
            raise ValueError("grad_slot/grad_coeff need one entry per element")
        ip = C.POINTER(C.c_int32)
        _check(self.lib.atx_adjoint(self.h, states_ptr, int(nturns), int(turn), int(npart),
                                    float(energy), float(rest_energy), lam_ptr,
                                    slots.ctypes.data_as(ip), coeffs.ctypes.data_as(ip),
                                    grad_ptr, int(nparams), stream), self.lib)

    def adjoint_batch(self, items, energy: float, rest_energy: float, slots: np.ndarray,
                      coeffs: np.ndarray, grad_ptr: int, nparams: int,
                      stream: int | None = None) -> None:
        """Independent blocks of turns in one launch (atx_adjoint_batch).
        items: sequence of (states_ptr, lam_ptr, nturns, turn, npart,
        accumulate); grad_ptr may be None when no item accumulates."""
        slots = np.ascontiguousarray(slots, np.int32)
        coeffs = np.ascontiguousarray(coeffs, np.int32)
        if slots.size != self.nelem or coeffs.size != self.nelem:
            raise ValueError("grad_slot/grad_coeff need one entry per element")
        arr = (AdjointItem * max(len(items), 1))()
        for i, (sp, lp, nt, t, n, acc) in enumerate(items):
            arr[i] = AdjointItem(sp, lp, int(nt), int(t), int(n), 1 if acc else 0)
        ip = C.POINTER(C.c_int32)
        _check(self.lib.atx_adjoint_batch(self.h, arr, len(items), float(energy),
                                          float(rest_energy), slots.ctypes.data_as(ip),
                                          coeffs.ctypes.data_as(ip), grad_ptr, int(nparams),
                                          stream), self.lib)

    def diffusion_matrices(self, orbits: np.ndarray, energy: float,
                           rest_energy: float) -> np.ndarray:
        """Diffusion matrix of every element from its entrance orbit
        (atx_diffusion_matrices): orbits (6, nelem) -> (nelem, 6, 6)."""
        o = np.asfortranarray(orbits, dtype=np.float64)
        if o.shape != (6, self.nelem):
            raise ValueError("orbits must have shape (6, number of elements)")
        b = np.zeros(36 * self.nelem, dtype=np.float64)
        _check(self.lib.atx_diffusion_matrices(self.h, o.ctypes.data, b.ctypes.data,
                                               float(energy), float(rest_energy), None, 0),
               self.lib)
        return b.reshape(self.nelem, 6, 6).transpose(0, 2, 1).copy()

    def element_matrices(self, orbits: np.ndarray, step: float, energy: float,
                         rest_energy: float, strict: bool = False) -> np.ndarray:
        """6x6 matrix of every element around its entrance orbit by central
        differences of step (atx_element_matrices): orbits (6, nelem) ->
        (nelem, 6, 6)."""
        o = np.asfortranarray(orbits, dtype=np.float64)
        if o.shape != (6, self.nelem):
            raise ValueError("orbits must have shape (6, number of elements)")
        m = np.zeros(36 * self.nelem, dtype=np.float64)
        _check(self.lib.atx_element_matrices(self.h, o.ctypes.data, float(step), m.ctypes.data,
                                             float(energy), float(rest_energy), None,
                                             ATX_STRICT if strict else 0), self.lib)
        return m.reshape(self.nelem, 6, 6).transpose(0, 2, 1).copy()

    def check(self, energy: float, rest_energy: float) -> None:
        _check(self.lib.atx_lattice_check(self.h, float(energy), float(rest_energy)), self.lib)

    @property
    def turn_counter(self) -> int:
        t = C.c_int64()
        _check(self.lib.atx_lattice_turn_counter(self.h, C.byref(t)), self.lib)
        return int(t.value)

    def track(self, r_in: np.ndarray, nturns: int, refpts=None, turn: int = 0,
              energy: float = 0.0, rest_energy: float = 0.0, charge: float = -1.0,
              losses: bool = False, start_elem: int = 0, end_elem: int = 0,
              keep_counter: bool = False, strict: bool = False, repack: bool = False):
        """Host-array tracking with the at.c contract; returns rout or
        (rout, loss dict). ``repack``: as track_device."""
        if r_in.ndim not in (1, 2) or r_in.shape[0] != 6:
            raise ValueError("rin is not 6D")
        if r_in.dtype != np.float64:
            raise ValueError("rin is not a double array")
        if not r_in.flags.f_contiguous:
            raise ValueError("rin is not Fortran-aligned")
        npart = r_in.size // 6
        refs = np.zeros(0, np.uint32) if refpts is None else np.ascontiguousarray(refpts, np.uint32)
        rout = np.full((6, npart, refs.size, max(nturns, 0)), np.nan, order="F")
        flags = ((ATX_STRICT if strict else 0) | (ATX_KEEP_COUNTER if keep_counter else 0)
                 | (ATX_REPACK if repack else 0))
        lf = lt = le = lc = None
        if losses:
            lf = np.zeros(npart, np.uint8)
            lt = np.zeros(npart, np.uint32)
            le = np.zeros(npart, np.uint32)
            lc = np.zeros((6, npart), order="F")
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        rc = self.lib.atx_track(self.h, ptr(r_in), npart, int(nturns),
                                refs.ctypes.data_as(_u32p), refs.size, int(turn), float(energy),
                                float(rest_energy), float(charge), int(bool(losses)),
                                int(start_elem), int(end_elem), ptr(rout) if rout.size else None,
                                ptr(lf), ptr(lt), ptr(le), ptr(lc), None, flags)
        _check(rc, self.lib)
        if losses:
            return rout, {"islost": lf.astype(bool), "turn": lt, "elem": le, "coord": lc}
        return rout

    def track_device(self, r_ptr: int, npart: int, nturns: int, *, refpts=None,
                     rout_ptr: int | None = None, turn: int = 0, energy: float = 0.0,
                     rest_energy: float = 0.0, losses: bool = False, loss_ptrs=None,
                     stream: int | None = None, strict: bool = False, soa: bool = False,
                     start_elem: int = 0, end_elem: int = 0, repack: bool = False) -> None:
        """Asynchronous tracking of device-resident particles (raw device
        pointers, e.g. torch ``tensor.data_ptr()``) on a HIP stream.
        ``repack``: turn blocks with the survivors packed between them
        (ATX_REPACK; without refpts; identical results)."""
        refs = np.zeros(0, np.uint32) if refpts is None else np.ascontiguousarray(refpts, np.uint32)
        flags = (ATX_DEVICE_POINTERS | (ATX_STRICT if strict else 0) | (ATX_SOA if soa else 0)
                 | (ATX_REPACK if repack else 0))
        lf, lt, le, lc = loss_ptrs if loss_ptrs is not None else (None, None, None, None)
        rc = self.lib.atx_track(self.h, r_ptr, int(npart), int(nturns),
                                refs.ctypes.data_as(_u32p), refs.size, int(turn), float(energy),
                                float(rest_energy), -1.0, int(bool(losses)), int(start_elem),
                                int(end_elem), rout_ptr, lf, lt, le, lc, stream, flags)
        _check(rc, self.lib)

    def elempass(self, index: int, r_in: np.ndarray, energy: float = 0.0,
                 rest_energy: float = 0.0, strict: bool = False) -> np.ndarray:
        if r_in.shape[0] != 6 or r_in.dtype != np.float64 or not r_in.flags.f_contiguous:
            raise ValueError("rin must be a 6xN float64 Fortran-ordered array")
        rc = self.lib.atx_elempass(self.h, int(index), r_in.ctypes.data, r_in.size // 6,
                                   float(energy), float(rest_energy), None,
                                   ATX_STRICT if strict else 0)
        _check(rc, self.lib)
        return r_in


def track_batch(items: list, nturns: int, *, turn: int = 0, energy: float = 0.0,
                rest_energy: float = 0.0, losses: bool = True, block_size: int = 0,
                ppt: int = 0, unroll: int = 0, stream: int | None = None, strict: bool = False,
                soa: bool = False, continue_losses: bool = False, repack: bool = False,
                noalloc: bool = False) -> None:
    """One launch over many (DeviceLattice, device particle buffer) pairs.

    items: sequence of dicts with keys ``lattice`` (DeviceLattice), ``r``
    (device pointer), ``n`` (particles), with losses ``lost`` = (flag,
    turn, elem, coord) device pointers, and optionally ``nturns``/``turn``
    for an item that tracks its own number of turns. ``repack``: turn blocks
    with the survivors packed between them, run as one graph (ATX_REPACK).
    block_size, ppt, unroll: the launch shape (0: the engine's default).
    noalloc: through atx_track_batch_noalloc (no allocation; at most 64
    items, no repack, lattices already on the device)."""
    lib = load_library()
    arr = (BatchItem * max(len(items), 1))()
    for i, it in enumerate(items):
        b = arr[i]
        b.lattice = it["lattice"].h
        b.r = it["r"]
        b.num_particles = int(it["n"])
        b.nturns = int(it.get("nturns", -1))
        b.turn = int(it.get("turn", 0))
        if losses:
            b.lost_flag, b.lost_turn, b.lost_elem, b.lost_coord = it["lost"]
    flags = ((ATX_STRICT if strict else 0) | (ATX_SOA if soa else 0) | ATX_DEVICE_POINTERS
             | (ATX_CONTINUE_LOSSES if continue_losses else 0) | (ATX_REPACK if repack else 0))
    sh = launch_shape(block_size, ppt, unroll)
    fn = lib.atx_track_batch_noalloc if noalloc else lib.atx_track_batch
    _check(fn(arr, len(items), int(nturns), int(turn), float(energy), float(rest_energy),
              int(bool(losses)), None if sh is None else C.byref(sh), stream, flags), lib)


def device_count() -> int:
    lib = load_library()
    n = C.c_int(0)
    rc = lib.atx_device_count(C.byref(n))
    return int(n.value) if rc == ATX_OK else 0


def device_info(dev: int) -> list:
    lib = load_library()
    name = C.create_string_buffer(256)
    arch = C.create_string_buffer(64)
    plat = C.create_string_buffer(64)
    cus = C.c_int(0)
    _check(lib.atx_device_info(dev, name, 256, arch, 64, C.byref(cus), plat, 64), lib)
    return [name.value.decode(), arch.value.decode(), int(cus.value), plat.value.decode()]


def set_device(dev: int) -> None:
    lib = load_library()
    _check(lib.atx_set_device(int(dev)), lib)


def require_library() -> None:
    """Raise AtError if the engine is unusable."""
    try:
        load_library()
    except ImportError as exc:
        raise AtError(str(exc)) from exc
