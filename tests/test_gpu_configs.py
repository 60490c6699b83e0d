"""Parity at the workloads' own sizes.

The sweeps of the four GPU configurations run kernel paths that small shapes
never reach: 64x64 MFMA tiles (chosen once a grid has >= 1024 of them), the
split-K out-step of H_eff with its slice sum, the tiled MPO-bond GEMM of the
d=5 complex chain (K = D d^2 = 100), and the truncation split at 1024^2 and
1280^2.  Here each is compared with the numpy oracle or LAPACK at the shape
the configuration runs:

  C2  Heisenberg L=64,  chi=256, d=2, D=5, FP64
  C3  Josephson  L=32,  chi=256, d=5, D=4, complex128
  C4  Heisenberg L=128, chi=512, d=2, D=5, FP64 / FP32 / split-BF16
  C5  64 random-field Heisenberg chains, L 16-64, chi 32-256

and the whole sweeps are pinned to converged energies the reference holds
for the same Hamiltonians (its MI300X runs of dmrg2-gpu and quimb DMRG2):
  Heisenberg L=64,  chi=128:  -28.175424859530  (results.json, quimb-dmrg2)
  Heisenberg L=100, chi=512:  -44.127739893298
  Josephson  L=32,  chi=256, n_max=2, E_J=1, E_C=0.5, flux pi/4:  -12.840606887634
  (benchmarks/paper_results/mi300x/challenge/
   dmrg2-gpu_mi300x_challenge_20260407_104055.json).
"""
import numpy as np
import pytest

import bench
from conftest import rand, rel_err
from oracle import dmrg as od
from oracle import models

pytestmark = pytest.mark.gpu

# centre-bond shapes: (cL, cR, d, D, complex)
C4_BOND = (512, 512, 2, 5, False)
C2_BOND = (256, 256, 2, 5, False)
C3_BOND = (256, 256, 5, 4, True)


def _dev(dmrg_amd, a, dtype=None):
    return dmrg_amd.DeviceArray.from_numpy(a, dtype)


@pytest.mark.parametrize("shape", [C4_BOND, C2_BOND, C3_BOND], ids=["c4", "c2", "c3"])
def test_heff2_at_config_shape(dmrg_amd, ops, shape):
    """The two-site matvec at each configuration's centre bond against the
    oracle's tensordot contraction: C4 runs 64x64 tiles for X = L^T theta
    (2560 x 2048 x 512) and a four-slice split-K out-step (2048 x 512 x
    2560); C2 a two-slice split-K; C3 complex 64x64 tiles, the tiled
    MPO-bond GEMM (K = 100) and a two-slice complex split-K."""
    cL, cR, d, D, cplx = shape
    rng = np.random.default_rng(cL + 10 * d + D)
    L = rand(rng, (cL, D, cL), cplx)
    R = rand(rng, (cR, D, cR), cplx)
    if cplx:
        Ws = models.josephson_mpo(4, n_max=2, E_J=2.0, E_C=1.0, phi_ext=np.pi / 3)
    else:
        Ws = models.heisenberg_mpo(4)
    th = rand(rng, (cL, d, d, cR), cplx)
    ref = od.apply_heff2(L, R, Ws[1], Ws[2], th)
    out = dmrg_amd.DeviceArray(th.shape, th.dtype)
    ops.heff2(_dev(dmrg_amd, L), _dev(dmrg_amd, R), _dev(dmrg_amd, Ws[1]), _dev(dmrg_amd, Ws[2]),
              _dev(dmrg_amd, th), out)
    assert rel_err(out.numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16x3"])
def test_heff2_mixed_at_c4_shape(dmrg_amd, ops, prec):
    """The mixed modes' matvec at the C4 centre bond (FP32 MFMA with its
    two-slice split-K, or split-BF16 products) against the float64 contraction
    of the same float32 inputs: elementwise |out - ref| <= 5e-5 |L||theta||WW||R|
    (three FP32 stages of depth 512, 20 and 2560); dropping one bf16 term of
    an operand would leave ~4e-3."""
    cL, cR, d, D, _ = C4_BOND
    rng = np.random.default_rng(4 + prec)
    f32 = lambda a: a.astype(np.float32)
    L = f32(rng.uniform(-1, 1, (cL, D, cL)))
    R = f32(rng.uniform(-1, 1, (cR, D, cR)))
    th = f32(rng.uniform(-1, 1, (cL, d, d, cR)))
    Ws = models.heisenberg_mpo(4)
    W1, W2 = Ws[1], Ws[2]
    ref = od.apply_heff2(L.astype(np.float64), R.astype(np.float64), W1, W2, th.astype(np.float64))
    bound = od.apply_heff2(np.abs(L).astype(np.float64), np.abs(R).astype(np.float64), np.abs(W1),
                           np.abs(W2), np.abs(th).astype(np.float64))
    out = dmrg_amd.DeviceArray(th.shape, np.float32)
    ops.heff2_f32(_dev(dmrg_amd, L, np.float32), _dev(dmrg_amd, R, np.float32),
                  _dev(dmrg_amd, W1), _dev(dmrg_amd, W2), _dev(dmrg_amd, th, np.float32), out,
                  prec=prec)
    got = out.numpy().astype(np.float64)
    err = np.max(np.abs(got - ref) / bound)
    assert err < 5e-5, err


@pytest.mark.parametrize("shape", [(512, 512, 2, 5, False), (256, 256, 5, 4, True)],
                         ids=["c4", "c3"])
def test_env_updates_at_config_shape(dmrg_amd, ops, shape):
    """Left and right environment updates at the C4 and C3 centre bonds
    against the oracle's einsum-pinned contractions."""
    ci, co, d, D, cplx = shape
    rng = np.random.default_rng(ci + d)
    W = rand(rng, (D, d, d, D), cplx)
    Lg = rand(rng, (ci, D, ci), cplx)
    A = rand(rng, (ci, d, co), cplx)
    out = dmrg_amd.DeviceArray((co, D, co), A.dtype)
    ops.env_left(_dev(dmrg_amd, Lg), _dev(dmrg_amd, A), _dev(dmrg_amd, W), out)
    assert rel_err(out.numpy(), od.update_left_env(Lg, A, W)) < 1e-12
    Rg = rand(rng, (ci, D, ci), cplx)
    B = rand(rng, (co, d, ci), cplx)
    outR = dmrg_amd.DeviceArray((co, D, co), A.dtype)
    ops.env_right(_dev(dmrg_amd, Rg), _dev(dmrg_amd, B), _dev(dmrg_amd, W), outR)
    assert rel_err(outR.numpy(), od.update_right_env(Rg, B, W)) < 1e-12


@pytest.mark.parametrize("cplx,opA,opB", [(False, 0, 0), (False, 1, 0), (False, 0, 1), (False, 1, 1),
                                          (True, 3, 0), (True, 0, 3), (True, 2, 2), (True, 3, 3)])
def test_gemm_64x64_tiles(dmrg_amd, ops, cplx, opA, opB):
    """2048 x 2048 output: 1024 tiles of 64x64, the tile the large sweep GEMMs
    use, with every operand form (transposed, conjugated) and alpha/beta."""
    M = N = 2048
    K = 192
    rng = np.random.default_rng(opA * 4 + opB + (8 if cplx else 0))
    A = rand(rng, (K, M) if opA in (1, 3) else (M, K), cplx)
    B = rand(rng, (N, K) if opB in (1, 3) else (K, N), cplx)
    C0 = rand(rng, (M, N), cplx)
    f = {0: lambda x: x, 1: lambda x: x.T, 2: lambda x: x.conj(), 3: lambda x: x.conj().T}
    alpha, beta = (0.7 - 0.2j, 0.3 + 0.1j) if cplx else (0.7, 0.3)
    ref = alpha * f[opA](A) @ f[opB](B) + beta * C0
    dC = _dev(dmrg_amd, C0)
    ops.gemm(opA, opB, _dev(dmrg_amd, A), _dev(dmrg_amd, B), dC, alpha, beta)
    assert rel_err(dC.numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("opA,opB,M,N,K", [(1, 0, 2560, 2048, 512), (0, 0, 2048, 512, 2560)],
                         ids=["c4_x", "c4_out"])
def test_gemm_f32_at_c4_shapes(dmrg_amd, ops, prec, opA, opB, M, N, K):
    """The two GEMM shapes of the C4 mixed matvec, as one launch each (64x64
    FP32 tiles; split-BF16): |C - C_ref| <= 2e-6 (|A||B|) elementwise."""
    rng = np.random.default_rng(M + N + K + prec)
    A = rng.standard_normal((K, M) if opA else (M, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if opB else (K, N)).astype(np.float32)
    A64 = (A.T if opA else A).astype(np.float64)
    B64 = (B.T if opB else B).astype(np.float64)
    dC = dmrg_amd.DeviceArray((M, N), np.float32)
    ops.gemm_f32(opA, opB, _dev(dmrg_amd, A, np.float32), _dev(dmrg_amd, B, np.float32), dC,
                 prec=prec)
    err = np.max(np.abs(dC.numpy().astype(np.float64) - A64 @ B64) / (np.abs(A64) @ np.abs(B64)))
    assert err < 2e-6, err


def _graded(rng, m, n, cplx, decay):
    """m x n matrix U diag(s) V^H with s_j = exp(-j / decay): the graded
    spectrum of a DMRG two-site tensor."""
    def orth(k, r):
        q, _ = np.linalg.qr(rand(rng, (k, r), cplx))
        return q
    r = min(m, n)
    s = np.exp(-np.arange(r) / decay)
    return (orth(m, r) * s) @ orth(n, r).conj().T, s


@pytest.mark.parametrize("scale_left", [False, True], ids=["R", "L"])
@pytest.mark.parametrize("m,k,cplx,decay", [(1024, 512, False, 60.0), (1280, 256, True, 40.0)],
                         ids=["c4", "c3"])
def test_svd_split_at_config_shape(dmrg_amd, ops, scale_left, m, k, cplx, decay):
    """The truncated split of the C4 (1024^2, keep 512) and C3 (1280^2
    complex, keep 256) bonds against LAPACK on a graded spectrum: kept
    singular values to 1e-12 sigma_max, the truncation residual equal to the
    optimal one, the rank-k product equal to LAPACK's best rank-k
    approximation within the split's accuracy class, and the orthonormal
    factor orthonormal.

    The split diagonalises the Gram matrix rho = M M^H by Jacobi until every
    off-diagonal entry is below tau = 1e-14 ||rho||_F.  A coupling tau
    between the last kept and the first discarded direction turns the kept
    space by at most tau / (s_{k-1}^2 - s_k^2), which moves the rank-k product
    by that angle times ~s_{k-1}.  At the C4 cut (s = e^{-j/60}, k = 512) the
    Gram gap is 1.3e-9, so the product may differ from LAPACK's (which
    resolves the gap s_{k-1} - s_k directly) by ~1e-9 relative while its
    residual stays optimal to second order in the angle."""
    rng = np.random.default_rng(m + k + (1 if scale_left else 0))
    M, s = _graded(rng, m, m, cplx, decay)
    left, right, sig, nk = ops.svd_split(_dev(dmrg_amd, M), k, scale_left=scale_left)
    assert nk == k
    assert np.max(np.abs(sig - s[:k])) < 1e-12 * s[0]
    U, S, Vh = np.linalg.svd(M)
    best = (U[:, :k] * S[:k]) @ Vh[:k]
    tau = 1e-14 * np.sqrt(np.sum(S ** 4))
    angle = tau / (S[k - 1] ** 2 - S[k] ** 2)
    bound = 4 * angle * S[k - 1] / np.linalg.norm(best) + 1e-12
    assert rel_err(left @ right, best) < bound, bound
    resid = np.linalg.norm(M - left @ right)
    assert abs(resid - np.sqrt(np.sum(S[k:] ** 2))) < 1e-10 * np.linalg.norm(M)
    q = left if not scale_left else right.conj().T
    assert np.linalg.norm(q.conj().T @ q - np.eye(nk)) < 1e-12 * nk


# --- whole sweeps pinned to the reference's converged energies ------------

def _heisenberg_run(dmrg_amd, L, chi, sweeps, seed, **opt):
    eng = dmrg_amd.DMRG(L, 2, chi, 5, tol=1e-11, krylov_max=20, **opt)
    eng.set_mpo(bench.heisenberg_mpo(L))
    eng.initialize_mps_random(seed=seed, scale=0.1)
    E = eng.run(sweeps)
    e, n2 = eng.expectation()
    eng.close()
    return E, e, n2


@pytest.mark.parametrize("L,chi,sweeps,E_ref", [(64, 256, 10, -28.175424859530),
                                                (100, 512, 10, -44.127739893298)],
                         ids=["c2_L64_chi256", "L100_chi512"])
def test_heisenberg_pinned_to_reference(dmrg_amd, L, chi, sweeps, E_ref):
    """C2 itself (L=64, chi=256, seed 64) and the reference's L=100, chi=512
    run (the C4 bond dimension): the converged energy equals the reference's
    to 1e-10 per site, and <psi|H|psi> of the returned MPS equals it with
    norm 1."""
    E, e, n2 = _heisenberg_run(dmrg_amd, L, chi, sweeps, seed=L)
    assert abs(E - E_ref) < 1e-10 * L, (E, E_ref)
    assert abs(n2 - 1) < 1e-10 and abs(e - E) < 1e-10 * L


def test_josephson_pinned_to_reference(dmrg_amd):
    """The reference's complex Josephson run at the C3 size (L=32, d=5,
    chi=256; its harness's n_max=2, E_J=1, E_C=0.5, flux pi/4)."""
    L, chi = 32, 256
    Ws = models.josephson_mpo(L, n_max=2, E_J=1.0, E_C=0.5, phi_ext=np.pi / 4)
    eng = dmrg_amd.DMRG(L, 5, chi, 4, dtype="c128", tol=1e-11, krylov_max=20)
    eng.set_mpo(Ws)
    eng.initialize_mps_random(seed=32, scale=0.1)
    E = eng.run(15)
    e, n2 = eng.expectation()
    eng.close()
    assert abs(E - (-12.840606887634)) < 1e-10 * L, E
    assert abs(n2 - 1) < 1e-10 and abs(e - E) < 1e-10 * L


def test_c3_workload_consistent(dmrg_amd):
    """C3 as benchmarked (E_J=2, E_C=1, flux pi/3, 8 sweeps from the seed-32
    start; no reference energy exists for it): the returned MPS is
    normalised, <psi|H|psi> equals the returned energy, and the energy lies
    below the chi=64 run's (more states, lower variational energy)."""
    L = 32
    wl = bench.WORKLOADS["c3"]
    Ws = bench.workload_mpo(wl)
    out = {}
    for chi in (64, wl["chi"]):
        eng = dmrg_amd.DMRG(L, 5, chi, 4, dtype="c128", tol=0.0, krylov_max=wl["krylov"])
        eng.set_mpo(Ws)
        eng.initialize_mps_random(seed=wl["seed"], scale=0.1)
        E = eng.run(8)
        e, n2 = eng.expectation()
        eng.close()
        assert abs(n2 - 1) < 1e-10 and abs(e - E) < 1e-10 * L, (chi, E, e, n2)
        out[chi] = E
    assert out[wl["chi"]] < out[64] + 1e-10


def test_c3_rsvd_matches_exact_split(dmrg_amd):
    """C3 (complex Josephson L=32, d=5, chi=256) with the randomised split on
    the bonds where the reference would use it (m = 5 chi > 2 chi), two
    subspace iterations: the sweep energies equal the exact split's to
    1e-10 per site after 8 sweeps from the same seeded start."""
    wl = bench.WORKLOADS["c3"]
    Ws = bench.workload_mpo(wl)
    E = {}
    for rs in (0, 1):
        eng = dmrg_amd.DMRG(wl["L"], 5, wl["chi"], 4, dtype="c128", tol=0.0,
                            krylov_max=wl["krylov"], svd_rsvd=rs, rsvd_power=2)
        eng.set_mpo(Ws)
        eng.initialize_mps_random(seed=wl["seed"], scale=0.1)
        E[rs] = eng.run(8)
        if rs:
            assert eng.stats()["rsvd_calls"] > 0
        eng.close()
    assert abs(E[1] - E[0]) < 1e-10 * wl["L"], E


@pytest.fixture(scope="module")
def c4_fp64(dmrg_amd):
    wl = bench.WORKLOADS["c4"]
    return _heisenberg_run(dmrg_amd, wl["L"], wl["chi"], 10, wl["seed"])


@pytest.mark.parametrize("prec", [1, 2], ids=["fp32", "bf16x3"])
def test_c4_mixed_matches_fp64(dmrg_amd, c4_fp64, prec):
    """C4 (Heisenberg L=128, chi=512, seed 128), 10 sweeps in each mode from
    the same start: the mixed modes' energy agrees with the FP64 run's to
    |dE|/|E| <= 1e-10, and both are the <psi|H|psi> of their normalised MPS."""
    wl = bench.WORKLOADS["c4"]
    E64, e64, n64 = c4_fp64
    assert abs(n64 - 1) < 1e-10 and abs(e64 - E64) < 1e-10 * wl["L"]
    E, e, n2 = _heisenberg_run(dmrg_amd, wl["L"], wl["chi"], 10, wl["seed"], precision=prec)
    assert abs(E - E64) / abs(E64) <= 1e-10, (E, E64)
    assert abs(n2 - 1) < 1e-9 and abs(e - E) < 1e-9 * wl["L"]


C5_SWEEPS = 6


def test_c4_fp32_environments(dmrg_amd, c4_fp64):
    """C4 with the environment updates on FP32 MFMA as well (env_precision =
    1): the FP32 rounding that the environments accumulate along the chain
    bounds the agreement with the FP64 run at |dE|/|E| <= 5e-9.  The figure
    follows the rounding pattern, not a trend: 2.4e-10 and 1.1e-9 were
    measured with FP32 products that differ only in summation order (tile
    choices at the short edge bonds); with FP64 environments, the default, it
    is ~1e-11.  The energy is the FP64 <psi|H|psi> of the normalised state."""
    wl = bench.WORKLOADS["c4"]
    E64 = c4_fp64[0]
    E, e, n2 = _heisenberg_run(dmrg_amd, wl["L"], wl["chi"], 10, wl["seed"], precision=1,
                               env_precision=1)
    assert abs(E - E64) / abs(E64) <= 5e-9, (E, E64)
    assert abs(n2 - 1) < 1e-9 and abs(e - E) < 1e-9 * wl["L"]


@pytest.fixture(scope="module")
def c5_sequential(dmrg_amd):
    """Each of the 64 C5 chains solved alone in plain FP64 (6 sweeps, the
    benchmark's count: after 4 the larger chains, e.g. L=32 chi=256, are
    still moving by ~1e-9 relative between sweeps)."""
    chains = bench.batch_chains(bench.WORKLOADS["c5"])
    ref = []
    for i, (L, chi, W, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=20)
        e.set_mpo(bench.heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=1000 + i, scale=0.1)
        ref.append(e.run(C5_SWEEPS))
        e.close()
    return chains, np.array(ref)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32"])
def test_c5_batch_matches_sequential(dmrg_amd, c5_sequential, prec):
    """The 64 uneven chains of C5 run together by dmrg_run_batch: in FP64
    each energy equals its chain's own sequential solve exactly (the same
    kernels in the same order), in the FP32-matvec mode to |dE|/|E| <= 1e-10."""
    chains, ref = c5_sequential
    engines = []
    for i, (L, chi, W, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=20, precision=prec)
        e.set_mpo(bench.heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=1000 + i, scale=0.1)
        engines.append(e)
    E, T = dmrg_amd.run_batch(engines, C5_SWEEPS, threads=8)
    for e in engines:
        e.close()
    rel = np.abs(E - ref) / np.abs(ref)
    if prec == 0:
        assert np.array_equal(E, ref), np.max(rel)
    else:
        assert np.max(rel) <= 1e-10, np.max(rel)
    assert np.all(T > 0)
