"""Operator parity on the GPU: each C-ABI op against the numpy oracle.

Tolerances are relative Frobenius errors; FP64 accumulations of K <= 5120
terms stay below ~1e-14 relative, so 1e-12 leaves margin without hiding a
wrong index (which gives O(1) errors).
"""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import rand, rel_err
from oracle import dmrg as od
from oracle import models

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(dmrg_amd, a):
    return dmrg_amd.DeviceArray.from_numpy(a)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("opA,opB", [(0, 0), (1, 0), (0, 1), (1, 1), (3, 0), (0, 3), (2, 2), (3, 3)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (37, 53, 29), (130, 70, 260), (300, 257, 64)])
def test_gemm(dmrg_amd, ops, cplx, opA, opB, M, N, K):
    if not cplx and opA in (2, 3):
        opA -= 2
    if not cplx and opB in (2, 3):
        opB -= 2
    rng = np.random.default_rng(M * 1000 + N * 10 + K + 7 * opA + 3 * opB)
    A = rand(rng, (K, M) if opA in (1, 3) else (M, K), cplx)
    B = rand(rng, (N, K) if opB in (1, 3) else (K, N), cplx)
    C0 = rand(rng, (M, N), cplx)
    f = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj(), 3: lambda x: x.conj().T}
    alpha, beta = (0.7 - 0.2j, 0.3 + 0.1j) if cplx else (0.7, 0.3)
    ref = alpha * f[opA](A) @ f[opB](B) + beta * C0
    dA, dB, dC = _dev(dmrg_amd, A), _dev(dmrg_amd, B), _dev(dmrg_amd, C0)
    ops.gemm(opA, opB, dA, dB, dC, alpha, beta)
    got = dC.numpy()
    assert rel_err(got, ref) < TOL, (rel_err(got, ref), got.ravel()[:4], ref.ravel()[:4])


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("opA,opB,M,N,K", [(0, 0, 64, 64, 32), (1, 0, 200, 77, 45), (0, 0, 96, 130, 300),
                                            (1, 1, 33, 65, 129), (0, 1, 2048, 512, 320),
                                            (1, 0, 2560, 256, 512)])
def test_gemm_f32_mixed(dmrg_amd, ops, prec, opA, opB, M, N, K):
    """FP32 GEMM of the mixed matvec: FP32 MFMA (prec 1) and split-BF16
    products (prec 2, each operand as three bf16 terms, six cross products).
    Both must reach FP32-class accuracy against the float64 product of the
    same float32 inputs, |C - C_ref| <= 2e-6 (|A| |B|) elementwise; one bf16
    term per operand would leave ~4e-3."""
    rng = np.random.default_rng(M + 7 * N + 13 * K + opA + 2 * opB)
    A = rng.standard_normal((K, M) if opA else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if opB else (K, N)).astype(np.float32)
    f = (lambda x: x, lambda x: x.T)
    A64, B64 = f[opA](A).astype(np.float64), f[opB](B).astype(np.float64)
    ref = A64 @ B64
    scale = np.abs(A64) @ np.abs(B64)
    dA = dmrg_amd.DeviceArray.from_numpy(A, np.float32)
    dB = dmrg_amd.DeviceArray.from_numpy(B, np.float32)
    dC = dmrg_amd.DeviceArray((M, N), np.float32)
    ops.gemm_f32(opA, opB, dA, dB, dC, prec=prec)
    got = dC.numpy().astype(np.float64)
    err = np.max(np.abs(got - ref) / scale)
    assert err < 2e-6, err


def _rand_problem(rng, cL, cR, d, D, cplx):
    L = rand(rng, (cL, D, cL), cplx)
    R = rand(rng, (cR, D, cR), cplx)
    W1 = rand(rng, (D, d, d, D), cplx)
    W2 = rand(rng, (D, d, d, D), cplx)
    th = rand(rng, (cL, d, d, cR), cplx)
    return L, R, W1, W2, th


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("cL,cR,d,D", [(1, 2, 2, 5), (3, 4, 2, 5), (16, 16, 2, 5), (33, 20, 2, 5),
                                       (64, 64, 2, 5), (12, 9, 5, 4), (40, 40, 3, 4)])
def test_heff2(dmrg_amd, ops, cplx, cL, cR, d, D):
    rng = np.random.default_rng(cL * 97 + cR + d + D)
    L, R, W1, W2, th = _rand_problem(rng, cL, cR, d, D, cplx)
    ref = od.apply_heff2(L, R, W1, W2, th)
    if cL * cR <= 400:  # the reference's einsum identity (test_numerics.py:123), small shapes only
        ein = np.einsum("cwa,csre,wspm,mrqn,enf->apqf", L, th, W1, W2, R, optimize=True)
        assert rel_err(ref, ein) < 1e-13
    out = dmrg_amd.DeviceArray(th.shape, th.dtype)
    ops.heff2(_dev(dmrg_amd, L), _dev(dmrg_amd, R), _dev(dmrg_amd, W1), _dev(dmrg_amd, W2),
              _dev(dmrg_amd, th), out)
    assert rel_err(out.numpy(), ref) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("cL,cR,d,D", [(1, 2, 2, 5), (5, 7, 2, 5), (32, 32, 2, 5), (20, 25, 5, 4)])
def test_heff1(dmrg_amd, ops, cplx, cL, cR, d, D):
    rng = np.random.default_rng(cL * 31 + cR + d * D)
    L = rand(rng, (cL, D, cL), cplx)
    R = rand(rng, (cR, D, cR), cplx)
    W = rand(rng, (D, d, d, D), cplx)
    th = rand(rng, (cL, d, cR), cplx)
    ref = od.apply_heff1(L, R, W, th)
    out = dmrg_amd.DeviceArray(th.shape, th.dtype)
    ops.heff1(_dev(dmrg_amd, L), _dev(dmrg_amd, R), _dev(dmrg_amd, W), _dev(dmrg_amd, th), out)
    assert rel_err(out.numpy(), ref) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ci,co,d,D", [(1, 2, 2, 5), (3, 4, 2, 5), (30, 45, 2, 5), (64, 64, 2, 5),
                                       (10, 13, 5, 4)])
def test_env_updates(dmrg_amd, ops, cplx, ci, co, d, D):
    rng = np.random.default_rng(ci * 13 + co + d + D)
    Lg = rand(rng, (ci, D, ci), cplx)
    A = rand(rng, (ci, d, co), cplx)
    W = rand(rng, (D, d, d, D), cplx)
    refL = od.update_left_env(Lg, A, W)
    out = dmrg_amd.DeviceArray((co, D, co), A.dtype)
    ops.env_left(_dev(dmrg_amd, Lg), _dev(dmrg_amd, A), _dev(dmrg_amd, W), out)
    assert rel_err(out.numpy(), refL) < TOL
    Rg = rand(rng, (ci, D, ci), cplx)
    B = rand(rng, (co, d, ci), cplx)
    refR = od.update_right_env(Rg, B, W)
    outR = dmrg_amd.DeviceArray((co, D, co), A.dtype)
    ops.env_right(_dev(dmrg_amd, Rg), _dev(dmrg_amd, B), _dev(dmrg_amd, W), outR)
    assert rel_err(outR.numpy(), refR) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("scale_left", [False, True])
@pytest.mark.parametrize("m,n", [(4, 4), (1, 8), (8, 1), (30, 17), (17, 30), (64, 64), (130, 96),
                                 (256, 256), (8, 700), (700, 8)])
def test_svd_split(dmrg_amd, ops, cplx, scale_left, m, n):
    """Full-rank split against LAPACK: singular values to 1e-12 sigma_max, the
    orthonormal factor orthonormal to 1e-12, exact reconstruction, and the
    product factor's Gram diagonal (S^2) to 1e-12 sigma_max^2.  The very wide
    and tall shapes stage a k x max(m, n) product: the workspace is sized by
    max(m, n), not by the side of the density matrix."""
    rng = np.random.default_rng(m * 7 + n + (1000 if scale_left else 0))
    M = rand(rng, (m, n), cplx)
    k = min(m, n)
    left, right, sig, nk = ops.svd_split(_dev(dmrg_amd, M), k, scale_left=scale_left)
    s_ref = np.linalg.svd(M, compute_uv=False)
    assert nk == k
    assert np.max(np.abs(sig - s_ref[:nk])) < 1e-12 * s_ref[0]
    assert rel_err(left @ right, M) < 1e-12
    if not scale_left:  # left = U, right = S V^H
        orth, prod = left.conj().T @ left, right @ right.conj().T
    else:               # left = U S, right = V^H
        orth, prod = right @ right.conj().T, left.conj().T @ left
    # Frobenius norm of an O(eps) orthogonality defect over nk^2 entries
    assert np.linalg.norm(orth - np.eye(nk)) < 4e-14 * max(nk, 8)
    assert np.max(np.abs(prod - np.diag(sig ** 2))) < 1e-12 * s_ref[0] ** 2


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("scale_left", [False, True])
def test_svd_truncation_rank(dmrg_amd, ops, cplx, scale_left):
    """Rank-deficient input: the null space falls under the absolute 1e-14
    cutoff (rocsolver_dgesvdj semantics: sigma from the QR of A V), the kept
    sigma are accurate, and the truncated product reconstructs M.  A Gram
    (density-matrix) solver resolves directions to ~eps sigma_max^2 / sigma,
    so with sigma_min = 1e-4 the reconstruction bound is 1e-10."""
    rng = np.random.default_rng(5)
    m, n, r = 60, 48, 20
    U = np.linalg.qr(rand(rng, (m, m), cplx))[0][:, :r]
    V = np.linalg.qr(rand(rng, (n, n), cplx))[0][:, :r]
    s = np.logspace(0, -4, r)
    M = (U * s) @ V.conj().T
    left, right, sig, nk = ops.svd_split(_dev(dmrg_amd, M), 40, scale_left=scale_left)
    assert nk == r
    assert np.max(np.abs(sig - s) / s) < 1e-8
    assert rel_err(left @ right, M) < 1e-10


@pytest.mark.parametrize("cplx", [False, True])
def test_svd_relative_cutoff(dmrg_amd, ops, cplx):
    rng = np.random.default_rng(9)
    m, n = 40, 40
    U = np.linalg.qr(rand(rng, (m, m), cplx))[0]
    V = np.linalg.qr(rand(rng, (n, n), cplx))[0]
    s = np.logspace(0, -4, n)
    M = (U * s) @ V.conj().T
    left, right, sig, nk = ops.svd_split(_dev(dmrg_amd, M), 40, rel_cutoff=1e-2)
    assert nk == int(np.sum(s >= 1e-2))
    # discarded part is the optimal rank-nk error
    err = np.linalg.norm(M - left @ right)
    assert abs(err - np.linalg.norm(s[nk:])) < 1e-10


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("N", [1, 5, 64, 65, 200, 512])
def test_heig(dmrg_amd, ops, cplx, N):
    """Block-Jacobi Hermitian eigen-solver against LAPACK eigh on an
    indefinite matrix with a clustered and a degenerate part."""
    rng = np.random.default_rng(N + (7 if cplx else 0))
    X = rand(rng, (N, N), cplx)
    Q = np.linalg.qr(X)[0]
    lam = rng.uniform(-3, 3, N)
    lam[: N // 4] = np.round(lam[: N // 4], 1)  # repeated values
    H = (Q * lam) @ Q.conj().T
    H = (H + H.conj().T) / 2
    ev, vec, sweeps = ops.heig(_dev(dmrg_amd, H))
    ref = np.sort(np.linalg.eigvalsh(H))[::-1]
    scale = np.abs(lam).max()
    assert np.max(np.abs(ev - ref)) < 1e-12 * scale
    assert np.linalg.norm(vec.conj().T @ vec - np.eye(N)) < 1e-12 * max(1, np.sqrt(N))
    assert np.linalg.norm(H @ vec - vec * ev) < 1e-11 * scale * max(1, np.sqrt(N))


@pytest.mark.parametrize("k", [1, 2, 3, 7, 20, 64, 100])
def test_tridiag_lowest(ops, k):
    rng = np.random.default_rng(k)
    a = rng.uniform(-2, 2, k)
    b = rng.uniform(0.1, 1.0, max(k - 1, 0))
    lam, y = ops.tridiag_lowest(a, b)
    if k == 1:
        assert lam == a[0]
        return
    ev, evec = sla.eigh_tridiagonal(a, b)
    assert abs(lam - ev[0]) < 1e-14 * max(1, abs(ev[0]))
    assert min(np.linalg.norm(y - evec[:, 0]), np.linalg.norm(y + evec[:, 0])) < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
def test_lanczos2_matches_oracle(dmrg_amd, ops, cplx):
    rng = np.random.default_rng(11)
    cL, cR, d, D = 8, 8, 2, 5
    # Hermitian H_eff: environments from a random MPS of the Heisenberg chain
    Ws = models.heisenberg_mpo(6) if not cplx else models.josephson_mpo(6, n_max=1)
    d = Ws[0].shape[1]
    D = Ws[0].shape[0]
    mps, b = od.random_mps(6, d, 8, seed=3, complex_=cplx)
    dt = complex if cplx else float
    Lg = od.left_boundary(D, dt)
    for i in range(2):
        Lg = od.update_left_env(Lg, mps[i], Ws[i])
    Rg = od.right_boundary(D, dt)
    for i in range(5, 3, -1):
        Rg = od.update_right_env(Rg, mps[i], Ws[i])
    th = np.einsum("asm,mtb->astb", mps[2], mps[3])
    E_ref, psi_ref = od.lanczos(lambda x: od.apply_heff2(Lg, Rg, Ws[2], Ws[3], x), th, 2, 20)
    dth = _dev(dmrg_amd, th)
    E, nit = ops.lanczos2(_dev(dmrg_amd, Lg), _dev(dmrg_amd, Rg), _dev(dmrg_amd, Ws[2]),
                          _dev(dmrg_amd, Ws[3]), dth, krylov_max=20, site=2)
    assert abs(E - E_ref) < 1e-11 * max(1, abs(E_ref))
    psi = dth.numpy()
    ov = abs(np.vdot(psi_ref.ravel(), psi.ravel()))
    assert abs(ov - 1) < 1e-9


def _low_rank(rng, m, n, rank, cplx, decay):
    """m x n matrix of exact rank `rank` with singular values e^{-j/decay}."""
    def orth(k, r):
        q, _ = np.linalg.qr(rand(rng, (k, r), cplx))
        return q
    s = np.exp(-np.arange(rank) / decay)
    return (orth(m, rank) * s) @ orth(n, rank).conj().T, s


@pytest.mark.parametrize("scale_left", [False, True], ids=["R", "L"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_rsvd_split_exact_when_rank_fits_sketch(dmrg_amd, ops, cplx, scale_left):
    """The randomised split (sketch r = k + 10) of a matrix of rank k + 5 <= r:
    the sketch spans the whole range, so the result equals the exact
    truncated split -- kept product equal to LAPACK's best rank-k
    approximation and the orthonormal factor orthonormal.  m > 2k as on the
    bonds where the engine uses it."""
    m, n, k = 320, 240, 40
    rng = np.random.default_rng(11 + cplx + 2 * scale_left)
    M, s = _low_rank(rng, m, n, k + 5, cplx, 12.0)
    left, right, nk = ops.rsvd_split(dmrg_amd.DeviceArray.from_numpy(M), k, oversample=10,
                                     power=0, seed=77, scale_left=scale_left)
    assert nk == k
    U, S, Vh = np.linalg.svd(M)
    best = (U[:, :k] * S[:k]) @ Vh[:k]
    assert rel_err(left @ right, best) < 1e-10
    q = left if not scale_left else right.conj().T
    assert np.linalg.norm(q.conj().T @ q - np.eye(nk)) < 1e-12 * nk


@pytest.mark.parametrize("power,bound", [(0, 0.5), (2, 1e-5)])
def test_rsvd_split_graded_spectrum(dmrg_amd, ops, power, bound):
    """Full-rank graded spectrum (s_j = e^{-j/8}, m = 400, n = 300, k = 40):
    the randomised split's truncation residual |M - left right| exceeds the
    optimal sqrt(sum_{j>=k} s_j^2) by at most `bound` of it (two subspace
    iterations measured 2.3e-6), and the orthonormal factor stays
    orthonormal."""
    m, n, k = 400, 300, 40
    rng = np.random.default_rng(5)
    M, s = _low_rank(rng, m, n, n, False, 8.0)
    left, right, nk = ops.rsvd_split(dmrg_amd.DeviceArray.from_numpy(M), k, oversample=10,
                                     power=power, seed=5)
    assert nk == k
    opt = np.sqrt(np.sum(s[k:] ** 2))
    resid = np.linalg.norm(M - left @ right)
    assert opt <= resid * (1 + 1e-12)
    assert resid <= opt * (1 + bound), (resid / opt - 1)
    assert np.linalg.norm(left.T @ left - np.eye(nk)) < 1e-12 * nk


def test_rsvd_split_rank_deficient(dmrg_amd, ops):
    """Rank 12 < k = 20 < r: the dependent sketch columns are dropped and the
    cutoff keeps exactly the 12 nonzero singular values; the product
    reproduces M."""
    m, n, k = 200, 160, 20
    rng = np.random.default_rng(3)
    M, s = _low_rank(rng, m, n, 12, False, 4.0)
    left, right, nk = ops.rsvd_split(dmrg_amd.DeviceArray.from_numpy(M), k, seed=9)
    assert nk == 12
    assert rel_err(left @ right, M) < 1e-12
    assert np.linalg.norm(left.T @ left - np.eye(nk)) < 1e-12 * nk


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("cL,cR", [(5, 7), (16, 24), (33, 40), (64, 64), (128, 96)])
def test_heff2_f32_shapes(dmrg_amd, ops, prec, cL, cR):
    """The mixed modes' two-site matvec (FP32 MFMA, or split-BF16 products on
    operands split once into bf16 planes) at bond shapes whose sizes are and
    are not multiples of 8 (16-byte plane vectors, or element-wise loads at
    the tile edges): elementwise |out - ref| <= 5e-5 (|L| |theta| |WW| |R|)
    against the float64 contraction of the same float32 inputs."""
    d, D = 2, 5
    rng = np.random.default_rng(cL * 131 + cR + prec)
    f32 = lambda a: a.astype(np.float32)
    L = f32(rng.uniform(-1, 1, (cL, D, cL)))
    R = f32(rng.uniform(-1, 1, (cR, D, cR)))
    th = f32(rng.uniform(-1, 1, (cL, d, d, cR)))
    Ws = models.heisenberg_mpo(4)
    W1, W2 = Ws[1], Ws[2]
    ref = od.apply_heff2(L.astype(np.float64), R.astype(np.float64), W1, W2, th.astype(np.float64))
    bound = od.apply_heff2(np.abs(L).astype(np.float64), np.abs(R).astype(np.float64), np.abs(W1),
                           np.abs(W2), np.abs(th).astype(np.float64))
    dev = dmrg_amd.DeviceArray.from_numpy
    out = dmrg_amd.DeviceArray(th.shape, np.float32)
    ops.heff2_f32(dev(L, np.float32), dev(R, np.float32), dev(W1), dev(W2), dev(th, np.float32),
                  out, prec=prec)
    err = np.max(np.abs(out.numpy().astype(np.float64) - ref) / bound)
    assert err < 5e-5, err
