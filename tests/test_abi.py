"""CPU: the C ABI library loads, exports every declared symbol, and its host
logic (element registry, attribute validation, partial-turn checks) follows
the reference error semantics, all without a GPU (no compute calls)."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from at_amd import engine
from at_amd.lattice import Dipole, Drift, Element, Marker, Quadrupole, RFCavity
from at_amd.tracking import atpass, get_uint32_index, End

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "atx.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(atx_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(engine.EXPORTS)


def test_unknown_pass_method_is_runtime_error():
    with pytest.raises(RuntimeError, match="Not implemented PassMethod: UnknownPass"):
        engine.DeviceLattice([Drift("d", 1.0, PassMethod="UnknownPass")])


def test_missing_required_attribute_is_attribute_error():
    d = Drift("d", 1.0)
    del d.Length
    lat = engine.DeviceLattice([d])
    with pytest.raises(AttributeError, match="Length"):
        lat.check(0.0, 0.0)
    q = Quadrupole("q", 0.5, 1.0)
    del q.NumIntSteps
    with pytest.raises(AttributeError, match="NumIntSteps"):
        engine.DeviceLattice([q]).check(0.0, 0.0)
    b = Dipole("b", 1.0, 0.1)
    del b.ExitAngle
    with pytest.raises(AttributeError, match="ExitAngle"):
        engine.DeviceLattice([b]).check(0.0, 0.0)


def test_radiation_and_rf_need_energy():
    q = Quadrupole("q", 0.5, 1.0, PassMethod="StrMPoleSymplectic4RadPass")
    with pytest.raises(ValueError, match="Energy must be positive"):
        engine.DeviceLattice([q]).check(0.0, 0.0)
    engine.DeviceLattice([q]).check(6e9, 0.0)
    rf = RFCavity("rf", 0.0, 1e6, 3.5e8, 31, 6e9)
    with pytest.raises(ValueError):
        engine.DeviceLattice([rf]).check(0.0, 0.0)


def test_array_size_validation():
    d = Drift("d", 1.0)
    object.__setattr__(d, "T1", np.zeros(3))
    with pytest.raises(ValueError, match="T1"):
        engine.DeviceLattice([d]).check(0.0, 0.0)


def test_missing_pass_method_attribute_error():
    m = Marker("m")
    del m.PassMethod
    with pytest.raises(AttributeError):
        engine.DeviceLattice([m])


def test_partial_turn_errors_before_any_device_work():
    lat = engine.DeviceLattice([Drift("d", 1.0)] * 20)
    rin = np.asfortranarray(np.zeros((6, 1)))
    for kw, nt in ((dict(start_elem=30), 1), (dict(end_elem=30), 1),
                   (dict(start_elem=10, end_elem=5), 1), (dict(end_elem=10), 2)):
        with pytest.raises(ValueError):
            lat.track(rin, nt, None, **kw)


def test_atpass_argument_errors():
    # pyat/test/test_atpass.py:8-35
    rin = np.asfortranarray(np.zeros((6, 1)))
    with pytest.raises(TypeError):
        atpass(1, rin, 1)
    with pytest.raises(TypeError):
        atpass([], 1, 1)
    with pytest.raises(TypeError):
        atpass([], rin, "a")
    with pytest.raises(ValueError):
        atpass([], np.zeros((7, 1)), 1)
    with pytest.raises(ValueError):
        atpass([], np.zeros((2, 6)), 1)
    with pytest.raises(ValueError):
        atpass([], np.asfortranarray(np.zeros((2, 6))).T, 1)


def test_refpts_normalisation():
    lat = [Drift("d", 1.0)] * 5
    np.testing.assert_array_equal(get_uint32_index(lat, End), [5])
    np.testing.assert_array_equal(get_uint32_index(lat, None), [])
    np.testing.assert_array_equal(get_uint32_index(lat, [3, 0, -1, 3]), [0, 3, 5])
    np.testing.assert_array_equal(get_uint32_index(lat, [True, False, True, False, False, True]),
                                  [0, 2, 5])
    assert get_uint32_index(lat, 2).dtype == np.uint32


def test_integrator_types():
    # atgpu/SymplecticIntegrator.cpp:32-86: types 1..6, others rejected
    lat = engine.DeviceLattice([Quadrupole("q", 0.5, 1.0)])
    for t in range(1, 7):
        lat.set_integrator(t)
        lat.check(0.0, 0.0)
    with pytest.raises(RuntimeError, match="not supported"):
        lat.set_integrator(7)


def test_batch_argument_errors():
    import ctypes as C

    lib = engine.load_library()
    item = engine.BatchItem()
    item.lattice = None
    item.r = None
    item.num_particles = 10
    rc = lib.atx_track_batch(C.byref(item), 1, 5, 0, 3e9, 0.0, 0, None, None, 0)
    assert rc == engine.ATX_E_VALUE


def test_adjoint_batch_argument_errors():
    """atx_adjoint_batch rejects bad item lists before any device work."""
    lib = engine.load_library()
    lat = engine.DeviceLattice([Drift("d", 1.0)])
    slots = np.full(1, -1, np.int32)
    ip = C.POINTER(C.c_int32)
    sp, cp = slots.ctypes.data_as(ip), slots.ctypes.data_as(ip)

    def call(items, n=None, grad=None):
        arr = (engine.AdjointItem * max(len(items), 1))(*items)
        return lib.atx_adjoint_batch(lat.h, arr, len(items) if n is None else n, 3e9, 0.0, sp,
                                     cp, grad, 0, None)

    assert lib.atx_adjoint_batch(None, None, 0, 3e9, 0.0, None, None, None, 0, None) \
        == engine.ATX_E_VALUE
    assert call([], n=-1) == engine.ATX_E_VALUE
    assert call([], n=70000) == engine.ATX_E_VALUE
    # particles but NULL arrays
    assert call([engine.AdjointItem(None, None, 1, 0, 10, 0)]) == engine.ATX_E_VALUE
    # negative turn count
    assert call([engine.AdjointItem(8, 8, -1, 0, 10, 0)]) == engine.ATX_E_VALUE
    # accumulate without a gradient array
    assert call([engine.AdjointItem(8, 8, 1, 0, 10, 1)]) == engine.ATX_E_VALUE
    assert "grad is NULL" in lib.atx_last_error().decode()


def test_parallel_adjoint_sweep_cost_model():
    """The parallel sweep is chosen only while the beam leaves the device's
    SIMDs idle, with sub-blocks the shorter the fewer particles there are."""
    from at_amd.adjoint import sweep_plan, use_parallel_sweep

    simds = 1024
    assert sweep_plan(4096, [8] * 8, simds) == 1      # 64 waves on 1,024 SIMDs
    assert sweep_plan(8192, [8] * 8, simds) == 1
    assert sweep_plan(16384, [8] * 8, simds) == 0
    assert sweep_plan(128, [8] * 8, simds) == 1
    assert sweep_plan(4096, [8], simds) == 1          # one block still splits into turns
    assert sweep_plan(4096, [1], simds) == 0          # one turn: nothing to split
    assert sweep_plan(65536, [8] * 8, simds) == 0     # one wave per SIMD already
    assert sweep_plan(262144, [8] * 125, simds) == 0
    assert use_parallel_sweep(4096, 8, simds)
    assert not use_parallel_sweep(32768, 8, simds)


def test_device_count_without_gpu_is_zero_or_more():
    assert engine.device_count() >= 0


def test_invalid_launch_shapes_are_value_errors():
    """atx_lattice_set_launch_shape validates on the host: workgroup size a
    multiple of 64 in [64, 512], 1 or 2 particles per lane, 1, 2 or 4 slices
    per iteration (1 or 2 with two particles per lane)."""
    dl = engine.DeviceLattice([Drift("d", 1.0)])
    for shape in ((96, 1, 1), (1024, 1, 1), (256, 3, 1), (256, 1, 3), (256, 2, 4)):
        with pytest.raises(ValueError):
            dl.set_launch_shape(*shape)
    for shape in ((64, 1, 1), (512, 2, 2), (0, 2, 0), (0, 0, 0)):
        dl.set_launch_shape(*shape)


def test_element_matrices_argument_errors():
    """atx_element_matrices rejects a NULL lattice, NULL arrays and a zero
    step before any device work."""
    lib = engine.load_library()
    lat = engine.DeviceLattice([Drift("d", 1.0)])
    o = np.zeros(6)
    m = np.zeros(36)
    assert lib.atx_element_matrices(None, o.ctypes.data, 1e-8, m.ctypes.data, 3e9, 0.0, None,
                                    0) == engine.ATX_E_VALUE
    assert lib.atx_element_matrices(lat.h, None, 1e-8, None, 3e9, 0.0, None,
                                    0) == engine.ATX_E_VALUE
    assert lib.atx_element_matrices(lat.h, o.ctypes.data, 0.0, m.ctypes.data, 3e9, 0.0, None,
                                    0) == engine.ATX_E_VALUE
    assert "step" in lib.atx_last_error().decode()
    with pytest.raises(ValueError):
        lat.element_matrices(np.zeros((6, 2)), 1e-8, 3e9, 0.0)

