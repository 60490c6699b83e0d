"""CPU: pin the oracle's PRNG and MPO builders (no GPU).

* splitmix64 against the published reference outputs of the generator
  (Steele, Lea & Flood 2014; Vigna's splitmix64.c, seed 0).
* MPO builders against Kronecker-product Hamiltonians, the reference's own
  exact-diagonalisation known answers (gpu-rocm/dmrg2-gpu/src/
  test_dmrg2_gpu.cpp:208-211 Heisenberg, :292-297 Josephson) and the
  reference's benchmark MPO fixtures (benchmark_data/regular/*/mpo.npz,
  converted by tests/golden/make_fixtures.py).
"""
import os

import numpy as np
import pytest

from oracle import models, prng

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_splitmix64_published_vector():
    out = prng.splitmix64(0, 3)
    assert [int(x) for x in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_stream_range_and_offsets():
    a = prng.uniform_pm1(7, 1000)
    assert a.min() >= -1.0 and a.max() < 1.0
    # a stream split at an offset continues the same sequence
    b = np.concatenate([prng.uniform_pm1(7, 400), prng.uniform_pm1(7, 600, offset=400)])
    assert np.array_equal(a, b)


@pytest.mark.parametrize("L", [2, 3, 4, 6])
def test_heisenberg_mpo_equals_kronecker(L):
    H = models.mpo_to_dense(models.heisenberg_mpo(L))
    Hk = models.dense_heisenberg(L).toarray()
    assert np.abs(H - Hk).max() == 0.0


@pytest.mark.parametrize("L", [3, 4])
def test_josephson_mpo_equals_kronecker(L):
    kw = dict(n_max=1, E_J=1.3, E_C=0.7, phi_ext=0.4, n_g=0.1)
    H = models.mpo_to_dense(models.josephson_mpo(L, **kw))
    Hk = models.dense_josephson(L, **kw)
    Hk = Hk.toarray() if hasattr(Hk, "toarray") else Hk
    assert np.abs(H - Hk).max() < 1e-14
    assert np.abs(H - H.conj().T).max() == 0.0


def test_random_field_mpo_equals_kronecker():
    h = np.array([0.3, -1.2, 0.8, 0.05, -0.4])
    H = models.mpo_to_dense(models.random_field_heisenberg_mpo(5, h))
    Hk = models.dense_heisenberg(5, h=h).toarray()
    assert np.abs(H - Hk).max() < 1e-15


# Known answers from the reference's GPU driver (test_dmrg2_gpu.cpp:208-211).
@pytest.mark.parametrize("L,E", [(4, -1.616025403784), (8, -3.374932598688)])
def test_heisenberg_ed_known_answers(L, E):
    assert abs(models.ground_energy(models.dense_heisenberg(L)) - E) < 1e-11


def test_heisenberg_ed_L12():
    # quimb DMRG2 at chi=128 reached -5.142090632836 (benchmarks/paper_results/
    # mi300x/results.json); exact diagonalisation in-repo.
    E = models.ground_energy(models.dense_heisenberg(12))
    assert abs(E - (-5.142090632836)) < 1e-11


# Josephson known answers, n_max=1, E_J=1, E_C=0.5, phi=pi/4 (:292-297).
@pytest.mark.parametrize("L,E", [(4, -1.053346829927396), (6, -1.748843818181493)])
def test_josephson_ed_known_answers(L, E):
    H = models.dense_josephson(L, n_max=1, E_J=1.0, E_C=0.5, phi_ext=np.pi / 4)
    assert abs(models.ground_energy(H) - E) < 1e-11


def _fixture(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_fixture_heisenberg_mpo_matches_builder():
    """The reference's quimb-built MPO (L=12), cut to a 4-site chain
    (left edge, two bulk tensors, right edge), is the same operator as the
    closed-form builder."""
    W = _fixture("heis_L12_D20.npz")["W"]
    H = models.mpo_to_dense([W[0], W[1], W[2], W[11]])
    assert np.abs(H - models.dense_heisenberg(4).toarray()).max() == 0.0
    Hfull = models.mpo_to_dense(list(W[:1]) + list(W[1:3]) + list(W[-1:]))
    assert np.abs(Hfull - models.mpo_to_dense(models.heisenberg_mpo(4))).max() == 0.0


def test_fixture_josephson_mpo_matches_builder():
    """The reference's Josephson MPO (n_max=2, E_J=1, E_C=0.5, flux pi/4),
    cut to 3 sites, equals the builder's operator entry by entry."""
    W = _fixture("josephson_L20_D50_nmax2.npz")["W"]
    H = models.mpo_to_dense([W[0], W[1], W[19]])
    Ho = models.mpo_to_dense(models.josephson_mpo(3, n_max=2, E_J=1.0, E_C=0.5,
                                                  phi_ext=np.pi / 4))
    assert np.abs(H - Ho).max() < 1e-15


def test_fixture_heisenberg_initial_mps_left_canonical():
    z = _fixture("heis_L12_D20.npz")
    for i in range(11):
        A = z["mps_%d" % i]
        cL, d, cR = A.shape
        M = A.reshape(cL * d, cR, order="F")
        assert np.abs(M.T @ M - np.eye(cR)).max() < 1e-13


# ---------------------------------------------------------------------------
# Frustrated chains and the two-leg ladder (the CLI's --j1j2, --j1j2j3 and
# --ladder models; gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:329-470).

@pytest.mark.parametrize("L,Js", [(5, [0.9, -0.37]), (6, [1.0, 0.4, 0.2]), (6, [0.7, -0.3, 0.55])])
def test_spin_chain_mpo_is_the_kronecker_hamiltonian(L, Js):
    H = models.mpo_to_dense(models.spin_chain_mpo(L, Js))
    assert np.abs(H - models.dense_spin_chain(L, Js).toarray()).max() < 1e-14
    assert models.spin_chain_mpo(L, Js)[0].shape[0] == 2 + 3 * len(Js) * (len(Js) + 1) // 2


def test_spin_chain_range_one_is_heisenberg():
    # J2 = J3 = 0 reduces to the nearest-neighbour chain and its MPO
    for L in (4, 7):
        H = models.dense_spin_chain(L, [1.0, 0.0, 0.0]).toarray()
        assert np.abs(H - models.dense_heisenberg(L).toarray()).max() == 0.0
        Hm = models.mpo_to_dense(models.spin_chain_mpo(L, [1.0]))
        assert np.abs(Hm - models.mpo_to_dense(models.heisenberg_mpo(L))).max() < 1e-15


@pytest.mark.parametrize("L", [4, 6, 8])
def test_j1j2_majumdar_ghosh_closed_form(L):
    """At J2 = J1/2 a product of nearest-neighbour singlets is an exact ground
    state with E = -3L/8 J1 (even L); the reference tabulates the same values
    (test_dmrg2_gpu.cpp:344-346)."""
    E = models.ground_energy(models.dense_spin_chain(L, [1.0, 0.5]))
    assert abs(E - (-3.0 * L / 8.0)) < 1e-12
    # the dimer state itself: singlets on (0,1), (2,3), ...
    s = np.array([0.0, 1.0, -1.0, 0.0]) / np.sqrt(2.0)   # (|ud> - |du>)/sqrt 2
    psi = s
    for _ in range(L // 2 - 1):
        psi = np.kron(psi, s)
    H = models.dense_spin_chain(L, [1.0, 0.5])
    assert abs(psi @ (H @ psi) + 3.0 * L / 8.0) < 1e-12
    assert np.abs(H @ psi + 3.0 * L / 8.0 * psi).max() < 1e-12


@pytest.mark.parametrize("L,Js,E_ref", [(6, [1.0, 0.4, 0.2], -2.3229239173),
                                        (8, [1.0, 0.4, 0.2], -3.1205086380),
                                        (8, [1.0, 0.5, 0.25], -3.0690408638)])
def test_j1j2j3_known_answers(L, Js, E_ref):
    """test_dmrg2_gpu.cpp:390-394 (quoted to 10 decimals)."""
    assert abs(models.ground_energy(models.dense_spin_chain(L, Js)) - E_ref) < 1e-9


@pytest.mark.parametrize("L", [2, 3])
def test_ladder_mpo_is_the_kronecker_hamiltonian(L):
    H = models.mpo_to_dense(models.ladder_mpo(L, 0.8, 1.3))
    assert np.abs(H - models.dense_ladder(L, 0.8, 1.3).toarray()).max() < 1e-14


@pytest.mark.parametrize("L,E_ref", [(2, -2.0), (3, -3.1293852415718166), (4, -4.2930664566945656)])
def test_ladder_known_answers(L, E_ref):
    """test_dmrg2_gpu.cpp:442-447; L = 2 rungs is the 4-spin ring, E = -2."""
    assert abs(models.ground_energy(models.dense_ladder(L)) - E_ref) < 1e-9


def test_ladder_decoupled_rungs():
    # J_leg = 0: independent rung singlets, -3/4 J_rung each
    for L in (2, 3):
        E = models.ground_energy(models.dense_ladder(L, J_leg=0.0, J_rung=1.7))
        assert abs(E - (-0.75 * 1.7 * L)) < 1e-12
        Em = models.ground_energy(models.mpo_to_dense(models.ladder_mpo(L, 0.0, 1.7)))
        assert abs(Em - E) < 1e-12


def _tfim_free_fermion_energy(L, J, h):
    """Open transverse-field Ising chain H = -J sum sz_i sz_{i+1} - h sum sx_i
    (Pauli): by the Jordan-Wigner mapping to free fermions (Lieb, Schultz and
    Mattis 1961; Pfeuty 1970) the ground energy is minus the sum of the
    singular values of the L x L bidiagonal matrix with h on the diagonal and
    J above it."""
    M = np.diag(np.full(L, float(h))) + np.diag(np.full(L - 1, float(J)), 1)
    return -np.linalg.svd(M, compute_uv=False).sum()


@pytest.mark.parametrize("L", [2, 3, 5, 8, 10])
@pytest.mark.parametrize("J,h", [(1.0, 1.0), (1.0, 0.3), (0.7, 1.9)])
def test_tfim_ed_free_fermion_closed_form(L, J, h):
    """Pins dense_tfim (Pauli operators, both signs, open ends): spin-1/2
    operators in place of Pauli ones, or a missing bond, change E0."""
    E = models.ground_energy(models.dense_tfim(L, J=J, hx=h))
    assert abs(E - _tfim_free_fermion_energy(L, J, h)) < 1e-12 * L


@pytest.mark.parametrize("L", [2, 3, 5])
def test_tfim_mpo_equals_kronecker(L):
    H = models.mpo_to_dense(models.tfim_mpo(L, J=0.8, hx=0.6))
    assert np.abs(H - models.dense_tfim(L, J=0.8, hx=0.6).toarray()).max() < 1e-15


@pytest.mark.parametrize("L", [2, 3, 6])
def test_dense_heisenberg_closed_forms(L):
    """The all-up state is an exact eigenstate with E = (L-1)/4 (the top of the
    spectrum); two sites have the singlet ground energy -3/4."""
    H = models.dense_heisenberg(L).toarray()
    up = np.zeros(2 ** L)
    up[0] = 1.0
    assert np.abs(H @ up - (L - 1) / 4 * up).max() == 0.0
    assert np.linalg.eigvalsh(H)[-1] == pytest.approx((L - 1) / 4, abs=1e-13)
    if L == 2:
        assert models.ground_energy(models.dense_heisenberg(2)) == pytest.approx(-0.75, abs=1e-14)
