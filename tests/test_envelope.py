"""CPU: the equilibrium envelope (ohmi_envelope) and the normal-mode
analysis (a_matrix, get_tunes_damp) it reports with.

ohmi_envelope runs with the oracle doing the tracking and the reference's
own radiating cores giving the per-element diffusion matrices
(tests/oracle_hooks.py), against the MATLAB results the reference holds
for the hmba lattice (pyat/test/test_physics.py:438-580). The same code
runs on the engine in tests/test_gpu_envelope.py. a_matrix is pinned by
closed forms: block rotations with known tunes and damping, and the
symplecticity of A for a symplectic map."""

import os

import numpy as np
import pytest

from at_amd import envelope, load_mat
from at_amd.lattice import AtError
from oracle import pyoracle as po

import oracle_hooks

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# pyat/test/test_physics.py:445-580 (MATLAB results, observation point 121)
TUNES = np.array([3.8156301859417241e-01, 8.5437639666919862e-01, 1.0906043886603917e-04])
DAMPING = np.array([1.0082255849735316e-05, 6.5786188107455773e-06, 9.6535944754221195e-06])
MODE_MATRICES = np.array([[[ 6.9000150766975468e+00, -2.6006243587486107e-05,  3.2653578639235223e-19,
   -3.7967745269216019e-19,  7.0332461709808689e-08, -1.9588952305233929e-06],
  [-2.6003127378572992e-05,  1.4492721966186267e-01, -1.9213610720373289e-20,
    2.0416359702093117e-20, -2.6206635396777689e-08, -2.5026896439390100e-04],
  [ 5.6444262224190266e-26, -8.8841529063567627e-27,  3.8488698134843282e-45,
   -4.3573005366593365e-45,  2.1817740491973322e-33,  1.5325293690413872e-29],
  [-1.9696012530780710e-26,  3.1000916669610156e-27, -1.3430486127116434e-45,
    1.5204635969825715e-45, -7.6132183004964948e-34, -5.3477034630266425e-30],
  [-2.7096903009017986e-08, -1.8307416373635866e-06,  2.4141125168786456e-25,
   -2.5639333062849341e-25,  3.3076648482142071e-13,  3.1614417843383483e-09],
  [-2.2432905706817794e-06, -2.5027456550832711e-04,  3.3071665955236401e-23,
   -3.5131068233223940e-23,  4.5232871220285155e-11,  4.3218970718496246e-07]],

 [[-2.2928294576283527e-40,  1.5052548450645591e-57, -2.9188149330766947e-19,
   -1.5140778203379946e-19, -3.5051801544373919e-37, -2.9766163106410513e-35],
  [-6.2160514262232196e-41,  3.7882279090924725e-58, -1.4553702864044294e-19,
   -1.7875951010571767e-20, -9.5028350345329565e-38, -8.0698544767557654e-36],
  [ 8.3548833738190147e-22, -4.7882541912711381e-39,  2.6446809160227356e+00,
    2.6974783759959333e-06,  1.2772590345573058e-18,  1.0846546846854328e-16],
  [ 3.4232225141711210e-22, -2.4394082787927380e-39,  2.6974783759959333e-06,
    3.7811744847886963e-01,  5.2332769805336290e-19,  4.4441246760564963e-17],
  [-3.8353443543983292e-38,  2.9389076758041900e-55,  4.6703386529315050e-17,
   -5.8660977848835226e-17, -5.8633113211900941e-35, -4.9791529519320920e-33],
  [-9.9107960119587158e-38,  6.8871348009966869e-55, -3.9791944217300570e-17,
   -9.5586285962631032e-17, -1.5151203409488728e-34, -1.2866476816442893e-32]],

 [[ 9.1129972881601156e-07, -1.7503747040171271e-10,  1.3713152284793508e-19,
   -7.5075423107737713e-21,  5.2753100634258230e-04, -2.4327856196424929e-05],
  [-1.5258771338022758e-10,  2.9256399010740371e-14, -1.7833250548974794e-23,
    1.8515034658507326e-24, -8.8574315741309644e-08, -2.9379401411635980e-08],
  [-6.2113700707272211e-20,  1.1937412750029960e-23, -1.0034900096364489e-32,
    4.3194611901458771e-34, -3.5923395227888331e-17,  6.1469644287517179e-18],
  [ 2.1674341717930857e-20, -4.1655151798392467e-24,  3.5016405610563869e-33,
   -1.5072596996559408e-34,  1.2535333347903046e-17, -2.1449600658093055e-18],
  [ 5.2771395924912238e-04, -1.0135988186593673e-07,  7.9357371970076673e-17,
   -4.3535442566604639e-18,  3.0548430047810055e-01, -1.3745076013796528e-02],
  [-6.5095371736554014e-05,  1.7571715133671204e-08, -5.1141897662730421e-16,
   -5.7612397975267779e-17, -1.3745076036246284e-02,  3.2741090968146191e+00]]])
MODE_EMITTANCES = np.array([1.3203595267442241e-10, 1.1573828900796540e-37, 2.8586919200065274e-06])
R66 = np.array([[ 9.1365311679111472e-10, -3.4825080173542211e-15,  6.9592739907066549e-25,
   3.2842455917050560e-25,  1.5077619241445821e-09, -1.2683061573113966e-15],
 [-3.4825080173566153e-15,  1.9135586746581878e-11, -1.1673711355982052e-28,
  -5.5030817148786995e-29, -2.5110841375848474e-13, -1.2870496457572302e-13],
 [ 6.9592739907066539e-25, -1.1673711355981774e-28,  1.0481996276492947e-37,
   1.2657948880467825e-37,  4.0299728048696993e-22, -4.9337080646554969e-23],
 [ 3.2842455917050565e-25, -5.5030817148784842e-29,  1.2657948880467825e-37,
   5.5172109613132283e-38,  1.9018302161174277e-22, -2.3925455405223881e-23],
 [ 1.5077619241445821e-09, -2.5110841375848878e-13,  4.0299728048696993e-22,
   1.9018302161174277e-22,  8.7312080470949738e-07,  9.7941378315054599e-10],
 [-1.2683061573110081e-15, -1.2870496457572292e-13, -4.9337080646554964e-23,
  -2.3925455405223884e-23,  9.7941378315054641e-10,  9.3578103418469929e-06]])
R44 = np.array([[ 9.1104941490365636e-10, -3.0489008671947536e-15,  0.0000000000000000e+00,
   0.0000000000000000e+00],
 [-3.0489008671971892e-15,  1.9135586672600989e-11,  0.0000000000000000e+00,
   0.0000000000000000e+00],
 [ 0.0000000000000000e+00,  0.0000000000000000e+00,  0.0000000000000000e+00,
   0.0000000000000000e+00],
 [ 0.0000000000000000e+00,  0.0000000000000000e+00,  0.0000000000000000e+00,
   0.0000000000000000e+00]])
M66 = np.array([[-7.3563108998454019e-01,  4.6737140210495083e+00,  0.0000000000000000e+00,
   0.0000000000000000e+00,  2.9985148734129178e-03, -6.2769167863546892e-07],
 [-9.8166216409474760e-02, -7.3566631849557673e-01,  0.0000000000000000e+00,
   0.0000000000000000e+00,  1.6901521608948762e-04, -3.5380663193032396e-08],
 [ 0.0000000000000000e+00,  0.0000000000000000e+00,  6.0980448553604349e-01,
  -2.0960292415936137e+00,  0.0000000000000000e+00,  0.0000000000000000e+00],
 [ 0.0000000000000000e+00,  0.0000000000000000e+00,  2.9967517970374508e-01,
   6.0980020978885463e-01,  0.0000000000000000e+00,  0.0000000000000000e+00],
 [ 1.2468793675147450e-06,  2.1544876952915747e-05,  0.0000000000000000e+00,
   0.0000000000000000e+00,  9.9998069084148322e-01, -2.0933014695165365e-04],
 [ 1.7009958256745489e-04,  2.9957945846259553e-03,  0.0000000000000000e+00,
   0.0000000000000000e+00,  2.2432585580767217e-03,  9.9999953041829404e-01]])
ORBIT6 = np.array([-2.6352679435773159e-09, -1.4584877344363697e-10,  0.0000000000000000e+00,
  0.0000000000000000e+00, -6.6969442207863711e-06, -5.8845211247215208e-02])
EMITXY = np.array([1.320358475286751e-10, 0.000000000000000e+00])
EMITXYZ = np.array([1.3222438678441061e-10, 0.0000000000000000e+00, 2.8584082872712468e-06])


@pytest.fixture
def hooks(monkeypatch):
    if not (po.available("oracle") and po.available("ref")):
        pytest.skip("oracle not built")
    oracle_hooks.install(monkeypatch)


@pytest.fixture(scope="module")
def hmba6():
    return load_mat(os.path.join(GOLDEN, "hmba.mat")).radiation_on(copy=True)


@pytest.mark.parametrize("refpts", ([121], [0, 40, 121]))
def test_ohmi_envelope_known_answers(hooks, hmba6, refpts):
    emit0, beamdata, emit = envelope.ohmi_envelope(hmba6, refpts)
    assert len(emit) == len(refpts)
    obs = emit[-1]
    np.testing.assert_allclose(beamdata["tunes"], TUNES, rtol=2e-6)
    np.testing.assert_allclose(beamdata["damping_rates"], DAMPING, rtol=2e-6)
    np.testing.assert_allclose(beamdata["mode_matrices"], MODE_MATRICES, rtol=0, atol=1e-8)
    np.testing.assert_allclose(beamdata["mode_emittances"][[0, 2]], MODE_EMITTANCES[[0, 2]],
                               rtol=1e-6)
    assert abs(beamdata["mode_emittances"][1]) < 1e-30
    # moments compared relative to their correlation scale sqrt(R_ii R_jj);
    # the vertical plane has no emittance and holds rounding noise (~1e-25)
    hz = [0, 1, 4, 5]
    r66, ref = obs["r66"][np.ix_(hz, hz)], R66[np.ix_(hz, hz)]
    scale = np.sqrt(np.outer(np.diag(ref), np.diag(ref)))
    assert np.abs((r66 - ref) / scale).max() < 1e-6
    assert np.abs(obs["r66"][2:4]).max() < 1e-20
    np.testing.assert_allclose(obs["r44"], R44, rtol=1e-6, atol=1e-20)
    np.testing.assert_allclose(obs["m66"], M66, rtol=0, atol=1e-9)
    np.testing.assert_allclose(obs["orbit6"], ORBIT6, rtol=1e-7, atol=1e-20)
    np.testing.assert_allclose(obs["emitXY"], EMITXY, rtol=1e-9, atol=1e-18)
    np.testing.assert_allclose(obs["emitXYZ"], EMITXYZ, rtol=1e-9, atol=1e-12)
    # the moments at the start satisfy R = M R M^T + B over one turn
    bbcum, _ = envelope.diffusion_matrices_along(hmba6, orbit=emit0["orbit6"])
    m = emit0["m66"]
    r = emit0["r66"]
    np.testing.assert_allclose(m @ r @ m.T + bbcum[-1], r, rtol=0, atol=1e-9 * np.abs(r).max())


def test_ohmi_envelope_needs_6d(hooks):
    ring = load_mat(os.path.join(GOLDEN, "hmba.mat")).radiation_off(copy=True)
    with pytest.raises(AtError, match="radiation"):
        envelope.ohmi_envelope(ring)


def _rot(mu, beta, alpha, damp=0.0):
    c, s = np.cos(mu), np.sin(mu)
    gamma = (1 + alpha ** 2) / beta
    return np.exp(-damp) * np.array([[c + alpha * s, beta * s], [-gamma * s, c - alpha * s]])


def test_a_matrix_block_rotations():
    from scipy.linalg import block_diag

    tunes = np.array([0.31, 0.17, 0.004])
    damp = np.array([2e-5, 1e-5, 3e-5])
    M = block_diag(_rot(2 * np.pi * tunes[0], 7.0, 0.3, damp[0]),
                   _rot(2 * np.pi * tunes[1], 2.5, -0.8, damp[1]),
                   _rot(2 * np.pi * tunes[2], 40.0, 0.0, damp[2]).T)
    td = envelope.get_tunes_damp(M)
    np.testing.assert_allclose(td["tunes"], tunes, rtol=1e-12)
    np.testing.assert_allclose(td["damping_rates"], damp, rtol=1e-9)
    A, lam = envelope.a_matrix(M)
    R = np.linalg.inv(A) @ M @ A
    for k in range(3):
        blk = R[2 * k:2 * k + 2, 2 * k:2 * k + 2]
        # a damped rotation: equal diagonal, opposite off-diagonal
        assert abs(blk[0, 0] - blk[1, 1]) < 1e-12 and abs(blk[0, 1] + blk[1, 0]) < 1e-12
        np.testing.assert_allclose(np.hypot(blk[0, 0], blk[0, 1]), np.exp(-damp[k]), rtol=1e-12)
    # off-diagonal blocks vanish
    mask = np.kron(np.eye(3), np.ones((2, 2))) == 0
    assert np.abs(R[mask]).max() < 1e-12
    # uncoupled modes: each mode matrix is the Twiss matrix
    # [[beta, -alpha], [-alpha, gamma]] of its own plane and zero elsewhere
    mm = td["mode_matrices"]
    for k, (beta, alpha) in enumerate(((7.0, 0.3), (2.5, -0.8))):
        twiss = np.array([[beta, -alpha], [-alpha, (1 + alpha ** 2) / beta]])
        np.testing.assert_allclose(mm[k][2 * k:2 * k + 2, 2 * k:2 * k + 2], twiss, rtol=1e-9)
    for k in range(3):
        out = np.ones((6, 6), bool)
        out[2 * k:2 * k + 2, 2 * k:2 * k + 2] = False
        assert np.abs(mm[k][out]).max() < 1e-9


def test_a_matrix_symplectic_for_symplectic_map():
    rng = np.random.default_rng(3)
    from scipy.linalg import block_diag, expm

    # a coupled map, symplectic for the (delta, ct) form jmatswap: a stable
    # block rotation conjugated by exp(Sw H), H symmetric
    Sw = envelope.jmatswap(3)
    H = rng.normal(0, 0.1, (6, 6))
    H = H + H.T
    C = expm(Sw @ H)
    R0 = block_diag(_rot(1.9, 1.0, 0.0), _rot(1.1, 1.0, 0.0), _rot(0.05, 1.0, 0.0))
    M = C @ R0 @ np.linalg.inv(C)
    np.testing.assert_allclose(M.T @ Sw @ M, Sw, atol=1e-12)
    A, _ = envelope.a_matrix(M)
    # normalised modes: a_k^T Sw b_k = 1, and distinct modes are
    # symplectically orthogonal
    np.testing.assert_allclose(A.T @ Sw @ A, envelope.jmat(3), atol=1e-10)
    td = envelope.get_tunes_damp(M)
    mu = np.sort(np.minimum(td["tunes"], 1 - td["tunes"]) * 2 * np.pi)
    np.testing.assert_allclose(mu, [0.05, 1.1, 1.9], rtol=1e-10)
    np.testing.assert_allclose(td["damping_rates"], 0.0, atol=1e-12)


def test_a_matrix_unstable():
    M = np.diag([2.0, 0.5, 1.0, 1.0, 1.0, 1.0])
    with pytest.raises(AtError, match="Unstable"):
        envelope.a_matrix(M)
