"""CPU: pin the oracle's DMRG primitives and sweeps (no GPU).

Checks follow the reference's own tests where they exist
(cpu/pdmrg/pdmrg/tests/test_numerics.py) and otherwise what the mathematics
fixes (dense spectra, exact diagonalisation, variational bounds).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dmrg as od
from oracle import models
from oracle import pdmrg_serial as ps

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _r(rng, shape, cplx=False):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


@pytest.mark.parametrize("cplx", [False, True])
def test_apply_heff2_matches_reference_einsum(cplx):
    """test_numerics.py:112-126: 'awc,csre,wmsp,mnrq,fne->apqf' in the
    cpu/pdmrg conventions (L[bra, mpo, ket], W[wl, wr, up, down])."""
    rng = np.random.default_rng(42)
    cL, cR, d, D = 3, 4, 2, 5
    Lp = _r(rng, (cL, D, cL), cplx)        # pdmrg L[a(bra), w, c(ket)]
    Rp = _r(rng, (cR, D, cR), cplx)
    WLp = _r(rng, (D, D, d, d))            # pdmrg W[w, m, s, p]
    WRp = _r(rng, (D, D, d, d))
    thp = _r(rng, (cL, d, d, cR), cplx)    # theta[c, s, r, e]
    ref = np.einsum("awc,csre,wmsp,mnrq,fne->apqf", Lp, thp, WLp, WRp, Rp, optimize=True)
    # engine conventions: env (ket, mpo, bra) = pdmrg transposed; W[w, s(ket), sp(bra), wp]
    Lg, Rg = Lp.transpose(2, 1, 0), Rp.transpose(2, 1, 0)
    W1, W2 = WLp.transpose(0, 2, 3, 1), WRp.transpose(0, 2, 3, 1)
    got = od.apply_heff2(Lg, Rg, W1, W2, thp)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # the restated pdmrg routine agrees with its own test
    got2 = ps.apply_heff(Lp, Rp, WLp, WRp, thp)
    assert np.abs(got2 - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("cplx", [False, True])
def test_update_envs_match_reference_einsums(cplx):
    """test_numerics.py:162-192 (real there; complex tensors here as well,
    where the einsums' conjugate of the bra tensor matters)."""
    rng = np.random.default_rng(42)
    chi, chi2, D, d = 3, 4, 5, 2
    Lp = _r(rng, (chi, D, chi), cplx)
    A = _r(rng, (chi, d, chi2), cplx)
    Wp = _r(rng, (D, D, d, d))
    refL = np.einsum("awc,atp,wmst,csq->pmq", Lp, A.conj(), Wp, A, optimize=True)
    W = Wp.transpose(0, 2, 3, 1)
    gotL = od.update_left_env(Lp.transpose(2, 1, 0), A, W)          # (ket, mpo, bra)
    assert np.abs(gotL.transpose(2, 1, 0) - refL).max() < 1e-12
    assert np.abs(ps.update_left_env(Lp, A, Wp) - refL).max() < 1e-12
    Rp = _r(rng, (chi, D, chi), cplx)
    B = _r(rng, (chi2, d, chi), cplx)
    refR = np.einsum("awc,pta,mwst,qsc->pmq", Rp, B.conj(), Wp, B, optimize=True)
    gotR = od.update_right_env(Rp.transpose(2, 1, 0), B, W)
    assert np.abs(gotR.transpose(2, 1, 0) - refR).max() < 1e-12
    assert np.abs(ps.update_right_env(Rp, B, Wp) - refR).max() < 1e-12


def test_heff2_spectrum_equals_dense_4site():
    """test_numerics.py:73-110: with identity-like right sites, H_eff on
    sites (0,1) of a 4-site chain has the spectrum of the dense H."""
    L, d = 4, 2
    Ws = models.heisenberg_mpo(L)
    A2 = np.zeros((4, 2, 2))
    for s in range(2):
        for b in range(2):
            A2[2 * s + b, s, b] = 1.0
    A3 = np.zeros((2, 2, 1))
    for s in range(2):
        A3[s, s, 0] = 1.0
    D = Ws[0].shape[0]
    R = od.right_boundary(D, float)
    R = od.update_right_env(R, A3, Ws[3])
    R = od.update_right_env(R, A2, Ws[2])
    Lb = od.left_boundary(D, float)
    n = 1 * d * d * 4
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j] = od.apply_heff2(Lb, R, Ws[0], Ws[1], e.reshape(1, d, d, 4, order="F")).ravel(order="F")
    assert np.abs(H - H.T).max() < 1e-14
    ev = np.sort(np.linalg.eigvalsh(H))
    ref = np.sort(np.linalg.eigvalsh(models.dense_heisenberg(4).toarray()))
    assert np.abs(ev - ref).max() < 1e-12


def test_product_state_energy():
    """test_numerics.py:212-227: all-up product state has E = (L-1)/4."""
    L = 8
    mps, _ = od.product_mps(L, 2)
    assert abs(od.expectation(mps, models.heisenberg_mpo(L)) - (L - 1) / 4) < 1e-14


def test_expectation_equals_dense():
    L = 6
    mps, _ = od.random_mps(L, 2, 8, seed=9)
    psi = mps[0]
    for A in mps[1:]:
        psi = np.tensordot(psi, A, axes=([psi.ndim - 1], [0]))
    psi = psi.reshape(-1, order="F")
    H = models.dense_heisenberg(L).toarray()
    # Fortran flattening puts site 0 fastest; the Kronecker builder puts site 0
    # slowest, so reverse the site order of the dense vector.
    psi = psi.reshape([2] * L, order="F").transpose(list(range(L - 1, -1, -1))).reshape(-1, order="F")
    E_dense = psi @ H @ psi / (psi @ psi)
    E = od.expectation(mps, models.heisenberg_mpo(L)) / od.norm2(mps)
    assert abs(E - E_dense) < 1e-12


def test_lanczos_lowest_eigenvalue_and_residual():
    rng = np.random.default_rng(0)
    n = 120
    X = rng.standard_normal((n, n))
    H = (X + X.T) / 2
    v0 = rng.standard_normal((n,))
    stats = {}
    E, v = od.lanczos(lambda x: H @ x, v0, 0, krylov_max=n, stats=stats)
    ev = np.linalg.eigvalsh(H)
    assert abs(E - ev[0]) < 1e-10
    assert np.linalg.norm(H @ v - E * v) < 1e-6
    assert abs(np.linalg.norm(v) - 1.0) < 1e-12


def test_lanczos_invariant_subspace_breakdown():
    # start vector inside a 2-dimensional invariant subspace: exact after 2 steps
    H = np.diag([1.0, 2.0, 3.0, 4.0, 5.0])
    v0 = np.array([1.0, 1.0, 0, 0, 0])
    E, v = od.lanczos(lambda x: H @ x, v0, 0, krylov_max=5, stats={})
    assert abs(E - 1.0) < 1e-14


def test_svd_split2_truncation_rule():
    """accurate_svd.py:82-110 keeps min(k, max_bond); the GPU engine also
    drops sigma < 1e-14 (dmrg2_gpu_impl.h:1276)."""
    rng = np.random.default_rng(1)
    cL, d, cR = 6, 2, 6
    U = np.linalg.qr(rng.standard_normal((cL * d, cL * d)))[0]
    V = np.linalg.qr(rng.standard_normal((d * cR, d * cR)))[0]
    s = np.array([1, 0.5, 0.1, 1e-3, 1e-8, 1e-13, 5e-15, 1e-16, 0, 0, 0, 0])
    M = (U * s) @ V.T
    th = M.reshape(cL, d, d, cR, order="F")
    Al, Ar, S, disc = od.svd_split2(th, 100, "R")
    assert Al.shape[2] == 6            # first sigma below 1e-14 is index 6
    Al, Ar, S, disc = od.svd_split2(th, 3, "L")
    assert Al.shape[2] == 3
    assert abs(disc - np.sum(s[3:] ** 2)) < 1e-15
    rec = np.tensordot(Al, Ar, axes=([2], [0])).reshape(cL * d, d * cR, order="F")
    assert np.linalg.norm(rec - (U[:, :3] * s[:3]) @ V[:, :3].T) < 1e-12
    # the cutoff is absolute: with sigma_0 = 1e-3 the first sigma below 1e-14
    # is 1e-16 (index 5); a cutoff relative to sigma_0 would keep 6
    th3 = (th * 1e-3).copy()
    Al, Ar, S, disc = od.svd_split2(th3, 100, "R")
    assert Al.shape[2] == 5


@pytest.mark.parametrize("L,E", [(4, -1.616025403784), (8, -3.374932598688)])
def test_dmrg2_known_answers(L, E):
    mps, b = od.neel_mps(L, 2, 32)
    Ef, _, _, _ = od.dmrg2(models.heisenberg_mpo(L), mps, b, 32, 20, krylov_max=20)
    assert abs(Ef - E) < 1e-10


def test_dmrg1_known_answer_random_start():
    L = 8
    mps, b = od.random_mps(L, 2, 32, seed=4, scale=0.1)
    Ef, _, _, _ = od.dmrg1(models.heisenberg_mpo(L), mps, b, 32, 30, krylov_max=20)
    assert abs(Ef - (-3.374932598688)) < 1e-10


def test_dmrg2_josephson_known_answer():
    L = 6
    Ws = models.josephson_mpo(L, n_max=1, E_J=1.0, E_C=0.5, phi_ext=np.pi / 4)
    mps, b = od.random_mps(L, 3, 64, seed=6, complex_=True, scale=0.1)
    Ef, _, _, _ = od.dmrg2(Ws, mps, b, 64, 20, krylov_max=40)
    assert abs(Ef - (-1.748843818181493)) < 1e-10


def test_dmrg2_golden_L12_from_reference_fixture():
    """benchmark_data/regular/heisenberg/L12_D20: the reference's MPO and
    initial MPS, chi=20; quimb DMRG2 golden energy -5.142090628178122 and the
    exact ground energy -5.142090632840533 (chi=20 truncates)."""
    z = np.load(os.path.join(GOLDEN, "heis_L12_D20.npz"), allow_pickle=False)
    L = 12
    Ws = [z["W"][i] for i in range(L)]
    mps = [z["mps_%d" % i] for i in range(L)]
    bonds = [1] + [A.shape[2] for A in mps]
    rec = {}
    Ef, _, _, sweeps = od.dmrg2(Ws, mps, bonds, int(z["chi"]), 20, tol=1e-12, krylov_max=40,
                                record=rec)
    E_ed = -5.142090632840533
    assert Ef >= E_ed - 1e-12                       # variational
    assert abs(Ef - float(z["E_dmrg2"])) < 5e-9     # same truncation-limited optimum


@pytest.mark.parametrize("L", [10])
def test_pdmrg_serial_matches_oracle(L):
    """The restated cpu/pdmrg serial two-site sweep (the timed CPU baseline)
    reaches the oracle's / ED energy."""
    Ws = models.heisenberg_mpo(L)
    mps, _ = od.random_mps(L, 2, 32, seed=L, scale=0.1)
    E, _ = ps.serial_dmrg([np.array(A) for A in mps], [ps.to_pdmrg_mpo(W) for W in Ws], 32, 6)
    E_ed = models.ground_energy(models.dense_heisenberg(L))
    assert abs(E - E_ed) < 1e-9


@pytest.mark.parametrize("cplx", [False, True])
def test_apply_heff1_is_heff2_with_an_identity_site(cplx):
    """Single-site H_eff equals the (einsum-pinned) two-site one whose second
    site is a d = 1 site carrying the MPO bond through unchanged."""
    rng = np.random.default_rng(8)
    cL, cR, d, D = 3, 4, 3, 4
    Lg, Rg = _r(rng, (cL, D, cL), cplx), _r(rng, (cR, D, cR), cplx)
    W = _r(rng, (D, d, d, D), cplx)
    th = _r(rng, (cL, d, cR), cplx)
    W_id = np.eye(D).reshape(D, 1, 1, D)
    ref = od.apply_heff2(Lg, Rg, W, W_id, th[:, :, None, :])[:, :, 0, :]
    got = od.apply_heff1(Lg, Rg, W, th)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("two_site", [False, True])
def test_heff_is_hermitian_for_a_hermitian_mpo(two_site):
    """Environments of a Hermitian MPO (Josephson, complex) on a random complex
    MPS make H_eff Hermitian: <x, H y> = <H x, y>."""
    L = 6
    Ws = models.josephson_mpo(L, n_max=1, E_J=1.3, E_C=0.7, phi_ext=0.4)
    D, d = Ws[0].shape[0], Ws[0].shape[1]
    mps, b = od.random_mps(L, d, 8, seed=12, complex_=True)
    Lg = od.left_boundary(D, complex)
    for i in range(2):
        Lg = od.update_left_env(Lg, mps[i], Ws[i])
    last = 3 if two_site else 2
    Rg = od.right_boundary(D, complex)
    for i in range(L - 1, last, -1):
        Rg = od.update_right_env(Rg, mps[i], Ws[i])
    rng = np.random.default_rng(2)
    shape = (b[2], d, d, b[4]) if two_site else (b[2], d, b[3])
    x, y = _r(rng, shape, True), _r(rng, shape, True)
    H = (lambda v: od.apply_heff2(Lg, Rg, Ws[2], Ws[3], v)) if two_site else \
        (lambda v: od.apply_heff1(Lg, Rg, Ws[2], v))
    lhs, rhs = np.vdot(x, H(y)), np.vdot(H(x), y)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("direction", ["R", "L"])
def test_svd_split2_factors(direction):
    """The orthonormal factor is an isometry, the kept singular values are
    numpy's, and the product is the best rank-k approximation."""
    rng = np.random.default_rng(4)
    cL, d, cR, k = 5, 2, 6, 4
    th = _r(rng, (cL, d, d, cR), True)
    M = th.reshape(cL * d, d * cR, order="F")
    s = np.linalg.svd(M, compute_uv=False)
    Al, Ar, S, disc = od.svd_split2(th, k, direction)
    assert np.abs(S - s[:k]).max() < 1e-13
    A = Al.reshape(cL * d, k, order="F")
    B = Ar.reshape(k, d * cR, order="F")
    Q = A if direction == "R" else B.conj().T
    assert np.abs(Q.conj().T @ Q - np.eye(k)).max() < 1e-13
    assert abs(np.linalg.norm(M - A @ B) ** 2 - np.sum(s[k:] ** 2)) < 1e-12
    assert abs(disc - np.sum(s[k:] ** 2)) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_to_pdmrg_env_converts_engine_layout(cplx):
    """The CPU baseline hands the engine's environments (ket, mpo, bra) to the
    restated pdmrg step through to_pdmrg_env; the converted environment must
    be the one the pdmrg update itself builds from the same tensors, pinned by
    the reference einsum (test_numerics.py:162-192): a plain reshape, or a
    transpose of the wrong axes, fails."""
    rng = np.random.default_rng(9)
    chi, chi2, D, d = 3, 4, 5, 2
    Lp = _r(rng, (chi, D, chi), cplx)
    A = _r(rng, (chi, d, chi2), cplx)
    Wp = _r(rng, (D, D, d, d))
    ref = np.einsum("awc,atp,wmst,csq->pmq", Lp, A.conj(), Wp, A, optimize=True)
    eng = od.update_left_env(Lp.transpose(2, 1, 0), A, Wp.transpose(0, 2, 3, 1))
    got = ps.to_pdmrg_env(eng)
    assert got.flags["C_CONTIGUOUS"]
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(np.ascontiguousarray(eng).reshape(ref.shape) - ref).max() > 1e-3
    assert np.abs(eng.transpose(0, 1, 2) - ref).max() > 1e-3
