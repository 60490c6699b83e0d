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
