"""CPU: the reference's own per-element tangent maps (diff_*.c), read off
its diffusion-matrix propagation by oracle/ref_driver.c:ref_tangent, pinned
against central finite differences of the oracle and against
symplecticity. These matrices are what the GPU adjoint is checked against.

Finding recorded here: for bend edges with FringeBendEntrance/Exit = 2 and
no gap correction, diff_bend_fringe.c:20-46 uses fy = irho*tan(edge)*pn for
the particle but leaves out d(fy)/d(delta) from the Jacobian, so the
delta column of its matrix differs from the derivative of the map in the
vertical rows; every other entry agrees."""

import numpy as np
import pytest

from at_amd import lattice as L
from oracle import pyoracle as po

pytestmark = pytest.mark.skipif(not po.available("ref"), reason="oracle/_ref not built")

J6 = np.zeros((6, 6))
for _i in range(3):
    J6[2 * _i, 2 * _i + 1] = 1.0
    J6[2 * _i + 1, 2 * _i] = -1.0


def fd_jacobian(elem, r6, h=1e-7):
    def f(r):
        a = np.asfortranarray(r.reshape(6, 1).copy())
        po.elempass(elem, a, impl="oracle")
        return a[:, 0]

    jac = np.zeros((6, 6))
    for j in range(6):
        d = np.zeros(6)
        d[j] = h
        jac[:, j] = (f(r6 + d) - f(r6 - d)) / (2 * h)
    return jac


def elements():
    return {
        "quad": L.Quadrupole("Q", 0.5, 1.3, NumIntSteps=10),
        "defoc_quad": L.Quadrupole("QD", 0.31, -2.1, NumIntSteps=7),
        "sext": L.Sextupole("S", 0.2, 35.0, NumIntSteps=8),
        "oct": L.Octupole("O", 0.1, [], [0.0, 0.0, 0.0, 500.0], NumIntSteps=4),
        "skew": L.Multipole("M", 0.3, [0.0, 0.4, 2.0], [0.0, 0.7, -3.0], NumIntSteps=6),
        "bend": L.Dipole("B", 1.0, 0.03, 0.0, EntranceAngle=0.015, ExitAngle=0.015),
        "cf_bend": L.Dipole("BQ", 1.0669994258, 0.0292, -0.9, EntranceAngle=0.0146,
                            ExitAngle=0.0146, NumIntSteps=10),
        "bend_m0": L.Dipole("B0", 0.8, 0.02, 0.2, EntranceAngle=0.01, ExitAngle=0.0,
                            FringeBendEntrance=0, FringeBendExit=0),
    }


POINTS = [np.array([1e-3, -2e-4, 5e-4, 1e-4, 2e-3, 0.0]),
          np.array([-3e-3, 4e-4, -1e-3, -2e-4, -0.02, 1e-3]),
          np.zeros(6)]


@pytest.mark.parametrize("name", list(elements()))
@pytest.mark.parametrize("ip", range(len(POINTS)))
def test_reference_tangent_equals_fd_of_oracle(name, ip):
    e = elements()[name]
    r6 = POINTS[ip]
    m = po.tangent_ref(e, r6)
    jac = fd_jacobian(e, r6)
    scale = max(1.0, np.abs(jac).max())
    # central FD error at h = 1e-7: O(h^2 f''') + O(eps/h) ~ 1e-9 at worst
    assert np.abs(m - jac).max() <= 2e-8 * scale, (m - jac)


@pytest.mark.parametrize("name", list(elements()))
def test_reference_tangent_is_symplectic(name):
    e = elements()[name]
    m = po.tangent_ref(e, POINTS[0])
    # in the [x px y py delta ct] ordering the pairs are (x, px), (y, py)
    # and (delta, ct), each with the 2x2 block [[0, 1], [-1, 0]]
    res = m.T @ J6 @ m - J6
    assert np.abs(res).max() < 1e-12, res


def test_method2_edge_omits_delta_dependence():
    e = L.Dipole("B2", 1.0, 0.05, 0.0, EntranceAngle=0.025, ExitAngle=0.025,
                 FringeBendEntrance=2, FringeBendExit=2)
    r6 = np.array([1e-3, -2e-4, 2e-3, 1e-4, 1e-2, 0.0])
    m = po.tangent_ref(e, r6)
    jac = fd_jacobian(e, r6)
    diff = np.abs(m - jac)
    # every column but d/d(delta) agrees, and so do the horizontal and
    # momentum rows of that column
    assert diff[:, [0, 1, 2, 3, 5]].max() < 1e-8
    assert diff[[0, 1, 4], 4].max() < 1e-8
    # the omitted term, y * irho*tan(edge)/(1+delta)^2 in d(py)/d(delta) at
    # each edge (5e-5 * 2e-3 ~ 1e-7 per edge), carried into y and py
    assert diff[3, 4] > 1e-6 and diff[2, 4] > 1e-6
