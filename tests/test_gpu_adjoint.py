"""GPU: the adjoint kernel (atx_adjoint) and the autograd wrapper.

* Per-element tangent maps: with lambda = e_i on particle i, one adjoint
  pass returns row i of the element's Jacobian. It must equal the
  reference's own diff_* Jacobian (read off its diffusion-matrix
  propagation, oracle/ref_driver.c:ref_tangent) and central finite
  differences of the oracle.
* Strength gradients of L = sum_alive (x^2+px^2+y^2+py^2) after 64 turns of
  the large ring (radiation off, cavity on), with respect to PolynomB[1] of
  every quadrupole and PolynomB[2] of every sextupole, against central
  finite differences (h = 1e-7*|K|) of the compiled reference tracking.
* Size-independent checks at the full 4,096-particle size: the gradient is
  additive over particle subsets, and independent of the checkpoint spacing.
* atx_adjoint_batch: one item equals atx_adjoint bit for bit (adjoint) and
  to rounding (gradient, accumulated with atomics); items of different
  sizes in one launch equal their own calls; the block-parallel sweep
  (unit seeds through the blocks, then one launch over all blocks) equals
  the sequential one.
"""

import copy
import os

import numpy as np
import pytest

from at_amd import engine
from at_amd import lattice as L
from at_amd.adjoint import Problem, loss_and_grad
from at_amd.workloads import gaussian_beam, large_ring
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ENERGY = 6.0e9
THREADS = min(16, len(os.sched_getaffinity(0)))


def _one(elem):
    return L.Lattice([elem], energy=ENERGY, particle="electron")


def _device_jacobian(elem, points):
    """Rows of the element Jacobian at each point, through atx_adjoint."""
    import torch

    lat = _one(elem)
    dl = engine.DeviceLattice(lat)
    npnt = len(points)
    n = 6 * npnt
    dev = torch.device("cuda:0")
    start = torch.tensor(np.repeat(np.asarray(points), 6, axis=0), dtype=torch.float64, device=dev)
    states = torch.empty((2, n, 6), dtype=torch.float64, device=dev)
    states[0].copy_(start)
    states[1].copy_(start)
    stream = torch.cuda.current_stream().cuda_stream
    rest = lat.particle.rest_energy
    dl.track_device(states[1].data_ptr(), n, 1, energy=ENERGY, rest_energy=rest, stream=stream)
    lam = torch.eye(6, dtype=torch.float64, device=dev).repeat(npnt, 1).contiguous()
    dl.adjoint(states.data_ptr(), 1, 0, n, ENERGY, rest, lam.data_ptr(),
               np.full(1, -1, np.int32), np.zeros(1, np.int32), 0, 0, stream=stream)
    torch.cuda.synchronize()
    return lam.cpu().numpy().reshape(npnt, 6, 6)


def _fd(elem, r6, h=1e-7):
    lat = _one(elem)

    def f(r):
        a = np.asfortranarray(r.reshape(6, 1).copy())
        po.atpass(lat, a, 1)
        return a[:, 0]

    jac = np.zeros((6, 6))
    for j in range(6):
        d = np.zeros(6)
        d[j] = h
        jac[:, j] = (f(r6 + d) - f(r6 - d)) / (2 * h)
    return jac


POINTS = [np.array([1e-3, -2e-4, 5e-4, 1e-4, 2e-3, 0.0]),
          np.array([-3e-3, 4e-4, -1e-3, -2e-4, -0.02, 1e-3]),
          np.zeros(6)]


def _multipoles():
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


@pytest.mark.skipif(not po.available("ref"), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", list(_multipoles()))
def test_tangent_maps_equal_reference_diff_maps(name):
    elem = _multipoles()[name]
    rows = _device_jacobian(elem, POINTS)
    for ip, r6 in enumerate(POINTS):
        m_ref = po.tangent_ref(elem, r6)
        scale = max(1.0, np.abs(m_ref).max())
        err = np.abs(rows[ip] - m_ref).max()
        print(f"{name} point {ip}: max |M_adj - M_ref| = {err:.3e}", flush=True)
        assert err <= 1e-12 * scale, rows[ip] - m_ref


def _others():
    return {
        "drift": L.Drift("D", 2.7),
        "thin": L.ThinMultipole("T", [0.0, 0.1], [0.0, 1.5, 20.0]),
        "corrector": L.Corrector("C", 0.1, [1e-4, -2e-4]),
        "rf": L.RFCavity("RF", 0.0, 6e6, 352e6 * 1.0, 992, ENERGY, TimeLag=0.05),
        "misaligned_quad": L.Quadrupole("QM", 0.4, 1.1, T1=np.array([1e-4, 0, -2e-4, 0, 0, 0]),
                                        T2=np.array([-1e-4, 0, 2e-4, 0, 0, 0]),
                                        R1=_rot(1e-3), R2=_rot(-1e-3)),
        "method2_bend": L.Dipole("B2", 1.0, 0.05, 0.0, EntranceAngle=0.025, ExitAngle=0.025,
                                 FringeBendEntrance=2, FringeBendExit=2),
        "kicked_bend": L.Dipole("BK", 1.0, 0.03, 0.1, KickAngle=np.array([1e-4, 5e-5])),
    }


def _rot(psi):
    c, s = np.cos(psi), np.sin(psi)
    r = np.eye(6)
    r[0, 0] = r[1, 1] = r[2, 2] = r[3, 3] = c
    r[0, 2] = r[1, 3] = s
    r[2, 0] = r[3, 1] = -s
    return r


@pytest.mark.parametrize("name", list(_others()) + list(_multipoles()))
def test_tangent_maps_equal_fd_of_oracle(name):
    elem = {**_others(), **_multipoles()}[name]
    rows = _device_jacobian(elem, POINTS)
    for ip, r6 in enumerate(POINTS):
        jac = _fd(elem, r6)
        scale = max(1.0, np.abs(jac).max())
        err = np.abs(rows[ip] - jac).max()
        print(f"{name} point {ip}: max |M_adj - M_fd| = {err:.3e}", flush=True)
        assert err <= 2e-8 * scale, rows[ip] - jac


def _ring_problem(checkpoint=8, parallel=None, sub_block=None):
    ring = large_ring(radiation=False, cavity=True)
    params = []
    for i, e in enumerate(ring):
        if isinstance(e, L.Quadrupole) or (e.PassMethod.startswith("StrMPole") and e.MaxOrder == 1):
            params.append((i, 1))
        elif e.PassMethod.startswith("StrMPole") and e.MaxOrder == 2:
            params.append((i, 2))
    return ring, Problem(ring, params, checkpoint=checkpoint, parallel=parallel,
                         sub_block=sub_block)


def _loss_cpu(ring, beam, nturns, edits, impl):
    lat = ring.copy()
    for idx, k, v in edits:
        e = copy.copy(lat[idx])
        pb = np.array(e.PolynomB, dtype=np.float64)
        pb[k] = v
        e.PolynomB = pb
        lat[idx] = e
    r = beam.copy(order="F")
    po.atpass(lat, r, nturns, impl=impl, nthreads=THREADS)
    alive = np.all(np.isfinite(r), axis=0)
    q = r[:4, alive]
    return float(np.sum(q * q))


@pytest.mark.skipif(not po.available("ref"), reason="oracle/_ref not built")
def test_strength_gradients_equal_fd_of_reference_tracking():
    ring, prob = _ring_problem()
    nq = sum(1 for _, k in prob.params if k == 1)
    ns = sum(1 for _, k in prob.params if k == 2)
    print(f"{len(ring)} elements, {nq} quadrupole and {ns} sextupole strengths", flush=True)
    beam = gaussian_beam(128, seed=11)
    nturns = 64
    loss, grad = loss_and_grad(prob, beam, nturns)
    theta = prob.initial_strengths()
    l_ref = _loss_cpu(ring, beam, nturns, [], "ref")
    print(f"loss gpu {loss!r} reference {l_ref!r}", flush=True)
    assert abs(loss - l_ref) <= 1e-9 * abs(l_ref)
    # FD on a spread of parameters: first/last quadrupole and sextupole of
    # several families, in different cells
    sext = [i for i, (_, k) in enumerate(prob.params) if k == 2]
    picks = sorted(set([0, 7, len(theta) // 2, len(theta) - 1, sext[0], sext[3], sext[-1]]))
    gmax = np.abs(grad).max()
    for i in picks:
        idx, k = prob.params[i]
        h = 1e-7 * abs(theta[i])
        lp = _loss_cpu(ring, beam, nturns, [(idx, k, theta[i] + h)], "ref")
        lm = _loss_cpu(ring, beam, nturns, [(idx, k, theta[i] - h)], "ref")
        g_fd = (lp - lm) / (2 * h)
        print(f"param {i} elem {idx} PolynomB[{k}]={theta[i]:.6g}: adjoint {grad[i]:.9e} "
              f"fd {g_fd:.9e}", flush=True)
        # FD rounding: eps*L/h relative to the strength ~ 1e-16*L/(1e-7*|K|)
        tol = 1e-6 * abs(g_fd) + 1e-15 * l_ref / h + 1e-9 * gmax
        assert abs(grad[i] - g_fd) <= tol


def test_gradient_additive_over_particles_and_checkpoint_independent():
    ring, prob8 = _ring_problem(checkpoint=8)
    beam = gaussian_beam(4096, seed=5)
    nturns = 64
    loss, grad = loss_and_grad(prob8, beam, nturns)
    parts = [loss_and_grad(prob8, np.asfortranarray(beam[:, s:s + 1024]), nturns)
             for s in range(0, 4096, 1024)]
    lsum = sum(p[0] for p in parts)
    gsum = sum(p[1] for p in parts)
    gmax = np.abs(grad).max()
    print(f"4096 particles: loss {loss!r}, max |grad| {gmax:.3e}", flush=True)
    assert abs(lsum - loss) <= 1e-12 * abs(loss)
    assert np.abs(gsum - grad).max() <= 1e-11 * gmax
    _, prob1 = _ring_problem(checkpoint=64)
    _, g1 = loss_and_grad(prob1, np.asfortranarray(beam[:, :512]), nturns)
    _, prob3 = _ring_problem(checkpoint=3)
    _, g3 = loss_and_grad(prob3, np.asfortranarray(beam[:, :512]), nturns)
    _, g8 = loss_and_grad(prob8, np.asfortranarray(beam[:, :512]), nturns)
    assert np.abs(g1 - g8).max() <= 1e-11 * np.abs(g8).max()
    assert np.abs(g3 - g8).max() <= 1e-11 * np.abs(g8).max()


def test_backward_uses_its_own_forward_strengths():
    """Two forwards at different strengths on one Problem (one device
    lattice), then the backward of the first: its recompute and sweep run
    with the first forward's strengths, so the gradient equals
    loss_and_grad at those strengths."""
    import torch

    from at_amd.adjoint import TrackingLoss

    ring, prob = _ring_problem(checkpoint=8)
    beam = gaussian_beam(512, seed=9)
    nturns = 16
    theta_a = prob.initial_strengths()
    theta_b = theta_a * (1.0 + 1e-3)
    la_ref, ga_ref = loss_and_grad(prob, beam, nturns, strengths=theta_a)
    parts = torch.from_numpy(np.ascontiguousarray(beam.T)).to(torch.device("cuda:0"))
    ta = torch.tensor(theta_a, dtype=torch.float64, requires_grad=True)
    tb = torch.tensor(theta_b, dtype=torch.float64, requires_grad=True)
    loss_a = TrackingLoss.apply(prob, ta, parts, nturns)
    loss_b = TrackingLoss.apply(prob, tb, parts, nturns)
    loss_a.backward()
    assert float(loss_a.detach().cpu()) == la_ref
    assert float(loss_b.detach().cpu()) != la_ref
    # the gradient sums per-wave atomics: equal to rounding (a backward at
    # theta_b's strengths would be off by about 1e-3 of max|g|)
    gmax = np.abs(ga_ref).max()
    assert np.abs(ta.grad.numpy() - ga_ref).max() <= 1e-12 * gmax
    loss_b.backward()
    _, gb_ref = loss_and_grad(prob, beam, nturns, strengths=theta_b)
    assert np.abs(tb.grad.numpy() - gb_ref).max() <= 1e-12 * np.abs(gb_ref).max()


def test_adjoint_rejects_radiating_lattice():
    ring = large_ring(radiation=True, cavity=True, ncells=1)
    idx = next(i for i, e in enumerate(ring) if isinstance(e, L.Quadrupole))
    prob = Problem(ring, [(idx, 1)])
    with pytest.raises(RuntimeError, match="adjoint not available"):
        loss_and_grad(prob, gaussian_beam(64, seed=1), 2)


def _block_states(prob, beam, nturns, turn0):
    """Turn-start states of one block and its end state, (nturns+1, N, 6)."""
    import torch

    r = torch.from_numpy(np.ascontiguousarray(beam.T)).to(torch.device("cuda", 0))
    n = r.shape[0]
    st = torch.empty((nturns + 1, n, 6), dtype=torch.float64, device=r.device)
    stream = torch.cuda.current_stream().cuda_stream
    for t in range(nturns):
        st[t].copy_(r)
        prob.dl.track_device(r.data_ptr(), n, 1, turn=turn0 + t, energy=prob.energy,
                             rest_energy=prob.rest, stream=stream)
    st[nturns].copy_(r)
    return st


def test_adjoint_batch_items_equal_single_calls():
    import torch

    ring, prob = _ring_problem()
    npar = len(prob.params)
    n, nturns, turn0 = 1000, 3, 5
    beam = gaussian_beam(n, seed=21)
    st = _block_states(prob, beam, nturns, turn0)
    assert torch.isnan(st[-1][:, 0]).sum().item() < n  # some particles survive
    gen = torch.Generator().manual_seed(3)
    lam0 = torch.randn((n, 6), generator=gen, dtype=torch.float64).to(st.device)
    stream = torch.cuda.current_stream().cuda_stream

    def single(lam, grad, states=st, npart=n):
        prob.dl.adjoint(states.data_ptr(), nturns, turn0, npart, prob.energy, prob.rest,
                        lam.data_ptr(), prob.slots, prob.coeffs, grad.data_ptr(), npar,
                        stream=stream)

    lam_a = lam0.clone()
    grad_a = torch.zeros(npar, dtype=torch.float64, device=st.device)
    single(lam_a, grad_a)
    # one item
    lam_b = lam0.clone()
    grad_b = torch.zeros_like(grad_a)
    prob.dl.adjoint_batch([(st.data_ptr(), lam_b.data_ptr(), nturns, turn0, n, True)],
                          prob.energy, prob.rest, prob.slots, prob.coeffs, grad_b.data_ptr(),
                          npar, stream=stream)
    assert torch.equal(lam_a, lam_b)
    gmax = grad_a.abs().max().item()
    assert gmax > 0
    assert (grad_a - grad_b).abs().max().item() <= 1e-13 * gmax
    # three items in one launch: the block again (accumulating), the block
    # with another adjoint (not accumulating), and the first 300 particles
    # of the block (states of their own, not accumulating)
    lam_c = lam0.clone()
    lam_d = (2.0 * lam0 + 1.0).contiguous()
    lam_e = lam0[:300].clone()
    st_e = st[:, :300].contiguous()
    grad_c = torch.zeros_like(grad_a)
    prob.dl.adjoint_batch([(st.data_ptr(), lam_c.data_ptr(), nturns, turn0, n, True),
                           (st.data_ptr(), lam_d.data_ptr(), nturns, turn0, n, False),
                           (st_e.data_ptr(), lam_e.data_ptr(), nturns, turn0, 300, False)],
                          prob.energy, prob.rest, prob.slots, prob.coeffs, grad_c.data_ptr(),
                          npar, stream=stream)
    lam_d_ref = (2.0 * lam0 + 1.0).contiguous()
    scratch = torch.zeros_like(grad_a)
    single(lam_d_ref, scratch)
    assert torch.equal(lam_c, lam_a)
    assert torch.equal(lam_d, lam_d_ref)
    assert torch.equal(lam_e, lam_a[:300])
    assert (grad_c - grad_a).abs().max().item() <= 1e-13 * gmax
    # no item accumulates: no gradient array needed, gradient slots or not
    lam_f = (2.0 * lam0 + 1.0).contiguous()
    prob.dl.adjoint_batch([(st.data_ptr(), lam_f.data_ptr(), nturns, turn0, n, False)],
                          prob.energy, prob.rest, prob.slots, prob.coeffs, None, npar,
                          stream=stream)
    assert torch.equal(lam_f, lam_d_ref)


@pytest.mark.parametrize("checkpoint,sub_block", [(8, None), (8, 2), (8, 3), (8, 8), (5, None)])
def test_parallel_sweep_equals_sequential(checkpoint, sub_block):
    """The parallel sweep over sub-blocks of 1 (the cost model's choice
    here), 2, 3 (uneven: 3+3+2 per block) and 8 turns, and over blocks of 5
    turns, equals the sequential sweep."""
    ring, pseq = _ring_problem(checkpoint=checkpoint, parallel=False)
    _, ppar = _ring_problem(checkpoint=checkpoint, parallel=True, sub_block=sub_block)
    beam = gaussian_beam(4096, seed=8)
    nturns = 64
    l_seq, g_seq = loss_and_grad(pseq, beam, nturns)
    l_par, g_par = loss_and_grad(ppar, beam, nturns)
    gmax = np.abs(g_seq).max()
    err = np.abs(g_par - g_seq).max()
    print(f"checkpoint {checkpoint}, sub-block {sub_block}: loss {l_seq!r}, "
          f"max |g_par - g_seq| = {err:.3e} "
          f"of max |g| {gmax:.3e}", flush=True)
    assert l_par == l_seq
    assert err <= 1e-11 * gmax
