"""Two-site and single-site DMRG restated from the reference GPU engines.

Oracle, test-only.  Everything here follows
gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h (two-site) and
gpu-rocm/dmrg-gpu/src/dmrg_gpu_impl.h (single-site) step by step, with numpy
in place of rocBLAS/rocSOLVER.

Tensor conventions (logical numpy arrays; Fortran-order flattening gives the
reference device layouts):

* MPS site tensor ``A[a, s, b]`` shape (cL, d, cR)      -> A[a + s*cL + b*cL*d]
* two-site ``theta[a, s1, s2, b]``                      -> (cL, d, d, cR) col-major
* MPO ``W[w, s, sp, wp]`` (see oracle.models)
* left environment ``Lg[a, w, a2]``  (ket, mpo, bra)     -> L[a + w*cL + a2*cL*D]
* right environment ``Rg[b, w, b2]`` (ket, mpo, bra)     -> R[b + w*cR + b2*cR*D]

The (ket, mpo, bra) column-major environment equals the ``cpu/pdmrg``
numpy ``L[bra, mpo, ket]`` in C order (environments/update.py:16-127).

Pinning (tests/test_oracle_dmrg.py):
* ``apply_heff2`` / ``apply_heff1``: the reference einsum
  'awc,csre,wmsp,mnrq,fne->apqf' (cpu/pdmrg/pdmrg/tests/test_numerics.py:123);
  the dense 4-site spectrum (:73-110); Hermiticity.
* ``update_left_env`` / ``update_right_env``: the reference einsums
  (test_numerics.py:174-175, :190-191); the product-state energy (L-1)/4
  (:212-227); full contraction = <psi|H|psi> from the dense ED matrix.
* ``lanczos``: lowest eigenvalue vs ``numpy.linalg.eigvalsh`` of the dense
  operator; breakdown on a 1-dimensional Krylov space; Ritz residual.
* ``svd_split*``: reconstruction, orthonormality, singular values vs
  ``numpy.linalg.svd``, the 1e-14 absolute cutoff rule.
* ``dmrg2`` / ``dmrg1``: exact diagonalisation (Heisenberg L=4,8 known answers
  from test_dmrg2_gpu.cpp:208-211, Josephson :292-297, and in-repo ED up to
  L=12), the quimb golden energy for L=12 (benchmark_data golden_results.json).
"""
import numpy as np
import scipy.linalg as sla

from . import prng

# Constants hard-coded in the reference.
KRYLOV_MAX_REF = 100          # dmrg2_gpu_impl.h:138  min(100, theta_size_max)
LANCZOS_BETA_TOL = 1e-12      # :975
LANCZOS_EIG_TOL = 1e-12       # :976
SVD_ABS_CUTOFF = 1e-14        # :1276
ZERO_NORM = 1e-14             # :983


def bond_dims_init(L, d, chi_max):
    """bond[i] = min(d^min(i, L-i), chi_max), bond[0] = bond[L] = 1
    (dmrg2_gpu_impl.h:70-76)."""
    b = [1] * (L + 1)
    for i in range(1, L):
        e = min(i, L - i)
        b[i] = chi_max if e * np.log(d) > np.log(chi_max) + 1e-12 else min(d ** e, chi_max)
    return b


def random_mps(L, d, chi_max, seed, complex_=False, scale=1.0, bonds=None):
    """Seeded random MPS with entries scale*U[-1,1), Fortran order, one
    stream over all sites (engine's ``dmrg_init_mps_random``).  Like the
    reference's initialize_mps_random (dmrg2_gpu_impl.h:382-393) it is not
    normalised or canonicalised."""
    if bonds is None:
        bonds = bond_dims_init(L, d, chi_max)
    sizes = [bonds[i] * d * bonds[i + 1] for i in range(L)]
    flat = prng.random_tensor_stream(seed, sizes, complex_, scale)
    return [f.reshape(bonds[i], d, bonds[i + 1], order='F') for i, f in enumerate(flat)], list(bonds)


def neel_mps(L, d, chi_max, bonds=None):
    """initialize_mps_neel (dmrg2_gpu_impl.h:415-429)."""
    if bonds is None:
        bonds = bond_dims_init(L, d, chi_max)
    out = []
    for i in range(L):
        cL, cR = bonds[i], bonds[i + 1]
        A = np.zeros((cL, d, cR))
        s = i % 2
        for a in range(min(cL, cR)):
            A[a, s, a] = 1.0
        out.append(A)
    return out, list(bonds)


def product_mps(L, d):
    """initialize_mps_product (dmrg2_gpu_impl.h:396-413): all bonds 1, |0>."""
    out = []
    for _ in range(L):
        A = np.zeros((1, d, 1))
        A[0, 0, 0] = 1.0
        out.append(A)
    return out, [1] * (L + 1)


# ---------------------------------------------------------------------------
# Contractions
# ---------------------------------------------------------------------------

def apply_heff2(Lg, Rg, W1, W2, theta):
    """result[a2,s1p,s2p,b2] = sum L[a,w,a2] th[a,s1,s2,b] W1[w,s1,s1p,m]
    W2[m,s2,s2p,n] R[b,n,b2]  (dmrg2_gpu_impl.h:606-780; effective_ham.py:21-81)."""
    X = np.tensordot(Lg, theta, axes=([0], [0]))            # w a2 s1 s2 b
    Y = np.tensordot(X, W1, axes=([0, 2], [0, 1]))          # a2 s2 b s1p m
    Z = np.tensordot(Y, W2, axes=([1, 4], [1, 0]))          # a2 b s1p s2p n
    return np.tensordot(Z, Rg, axes=([1, 4], [0, 1]))       # a2 s1p s2p b2


def apply_heff1(Lg, Rg, W, theta):
    """Single-site H_eff (dmrg_gpu_impl.h:538-705)."""
    X = np.tensordot(Lg, theta, axes=([0], [0]))            # w a2 s b
    Y = np.tensordot(X, W, axes=([0, 2], [0, 1]))           # a2 b sp wp
    return np.tensordot(Y, Rg, axes=([1, 3], [0, 1]))       # a2 sp b2


def update_left_env(Lg, A, W):
    """Lnew[b,w',b2] = sum L[a,w,a2] A[a,s,b] W[w,s,sp,w'] conj(A[a2,sp,b2])
    (dmrg2_gpu_impl.h:786-860; update.py:16-80)."""
    X = np.tensordot(Lg, A, axes=([0], [0]))                # w a2 s b
    Y = np.tensordot(X, W, axes=([0, 2], [0, 1]))           # a2 b sp w'
    return np.tensordot(Y, A.conj(), axes=([0, 2], [0, 1]))  # b w' b2


def update_right_env(Rg, A, W):
    """Rnew[a,w,a2] = sum A[a,s,b] W[w,s,sp,wp] R[b,wp,b2] conj(A[a2,sp,b2])
    (dmrg2_gpu_impl.h:866-934; update.py:83-127)."""
    X = np.tensordot(A, Rg, axes=([2], [0]))                 # a s wp b2
    Y = np.tensordot(X, W, axes=([1, 2], [1, 3]))            # a b2 w sp
    return np.tensordot(Y, A.conj(), axes=([1, 3], [2, 1]))  # a w a2


def left_boundary(D, dtype):
    Lg = np.zeros((1, D, 1), dtype)
    Lg[0, 0, 0] = 1.0
    return Lg


def right_boundary(D, dtype):
    Rg = np.zeros((1, D, 1), dtype)
    Rg[0, D - 1, 0] = 1.0
    return Rg


def expectation(mps, Ws):
    """<psi|H|psi> by left-environment contraction (unnormalised)."""
    D = Ws[0].shape[0]
    dt = np.result_type(mps[0], Ws[0])
    Lg = left_boundary(D, dt)
    for A, W in zip(mps, Ws):
        Lg = update_left_env(Lg, A, W)
    return Lg[0, D - 1, 0]


def norm2(mps):
    E = np.ones((1, 1))
    for A in mps:
        E = np.einsum('ab,asc,bsd->cd', E, A, A.conj())
    return E[0, 0].real


# ---------------------------------------------------------------------------
# Lanczos (dmrg2_gpu_impl.h:970-1160)
# ---------------------------------------------------------------------------

def lanczos(matvec, theta, site, krylov_max=KRYLOV_MAX_REF, stats=None):
    """Ground state of the effective Hamiltonian, reference semantics.

    * v0 = theta/|theta|; a zero-norm theta is replaced by a seeded random
      vector (engine: splitmix64 stream 42+site; reference: srand(42+site)).
    * each step: alpha = Re<v_i|w>; w -= alpha v_i + beta_{i-1} v_{i-1};
      one classical Gram-Schmidt pass against all v_0..v_i; beta = |w|;
      v_{i+1} = w/beta (1/beta := 0 for beta <= 1e-300).
    * every 3rd iteration from iter 6 on (iter >= 4 and iter % 3 == 0):
      breakdown if some beta_j < 1e-12 (Krylov dim j+1); converged if the
      lowest Ritz value moved by < 1e-12 (Krylov dim iter+1).
    * returns (E, psi) with psi the normalised Ritz vector V y.
    """
    shape = theta.shape
    n = theta.size
    max_iter = min(krylov_max, n)
    v = theta.ravel(order='F').copy()
    nrm = np.linalg.norm(v)
    if nrm < ZERO_NORM:
        v = prng.random_tensor_stream(42 + site, [n], np.iscomplexobj(v))[0].astype(v.dtype)
        nrm = np.linalg.norm(v)
    v = v / nrm
    V = np.zeros((n, max_iter), dtype=v.dtype)
    V[:, 0] = v
    alpha = np.zeros(max_iter)
    beta = np.zeros(max_iter)
    prev = 1e30
    niter = max_iter
    nmv = 0
    for it in range(max_iter):
        vi = V[:, it]
        nmv += 1
        w = matvec(vi.reshape(shape, order='F')).ravel(order='F')
        a = np.vdot(vi, w).real
        alpha[it] = a
        w = w - a * vi
        if it > 0:
            w = w - beta[it - 1] * V[:, it - 1]
            c = V[:, :it + 1].conj().T @ w
            w = w - V[:, :it + 1] @ c
        else:
            c0 = np.vdot(V[:, 0], w)
            w = w - c0 * V[:, 0]
        b = np.linalg.norm(w)
        beta[it] = b
        inv = 1.0 / b if b > 1e-300 else 0.0
        if it + 1 < max_iter:
            V[:, it + 1] = w * inv
        if it >= 4 and it % 3 == 0:
            ncheck = it + 1
            small = np.nonzero(beta[:ncheck] < LANCZOS_BETA_TOL)[0]
            if small.size:
                niter = int(small[0]) + 1
                break
            e = sla.eigh_tridiagonal(alpha[:ncheck], beta[:ncheck - 1],
                                     eigvals_only=True, select='i', select_range=(0, 0))[0]
            if abs(e - prev) < LANCZOS_EIG_TOL:
                niter = ncheck
                break
            prev = e
    if niter == 1:
        evals, evecs = np.array([alpha[0]]), np.ones((1, 1))
    else:
        evals, evecs = sla.eigh_tridiagonal(alpha[:niter], beta[:niter - 1])
    E = float(evals[0])
    psi = V[:, :niter] @ evecs[:, 0].astype(V.dtype)
    psi = psi / np.linalg.norm(psi)
    if stats is not None:
        stats['matvecs'] = stats.get('matvecs', 0) + nmv
        stats.setdefault('niter', []).append(niter)
    return E, psi.reshape(shape, order='F')


# ---------------------------------------------------------------------------
# SVD truncation (dmrg2_gpu_impl.h:1166-1343; dmrg_gpu_impl.h:1118-1350)
# ---------------------------------------------------------------------------

def truncation_rank(S, k, cutoff=SVD_ABS_CUTOFF, rel_cutoff=0.0):
    """new_k = first i < k with S[i] < cutoff (or S[i] < rel_cutoff*S[0]),
    else k; at least 1 (scalar_traits.h:497-505)."""
    thr = max(cutoff, rel_cutoff * (S[0] if len(S) else 0.0))
    new_k = k
    for i in range(k):
        if S[i] < thr:
            new_k = i
            break
    return max(new_k, 1)


def svd_split2(theta, chi_max, direction, cutoff=SVD_ABS_CUTOFF, rel_cutoff=0.0):
    """Two-site split: theta (cL,d,d,cR) -> A_left (cL,d,k), A_right (k,d,cR).
    'R': A_left = U, A_right = S Vh.  'L': A_left = U S, A_right = Vh."""
    cL, d1, d2, cR = theta.shape
    M = theta.reshape(cL * d1, d2 * cR, order='F')
    U, S, Vh = np.linalg.svd(M, full_matrices=False)
    k = min(len(S), chi_max)
    nk = truncation_rank(S, k, cutoff, rel_cutoff)
    U, S, Vh = U[:, :nk], S[:nk], Vh[:nk]
    if direction == 'R':
        Al, Ar = U, S[:, None] * Vh
    else:
        Al, Ar = U * S[None, :], Vh
    disc = float(np.sum(np.linalg.svd(M, compute_uv=False)[nk:] ** 2))
    return (Al.reshape(cL, d1, nk, order='F'), Ar.reshape(nk, d2, cR, order='F'),
            S, disc)


# ---------------------------------------------------------------------------
# Two-site sweep (dmrg2_gpu_impl.h:1349-1457)
# ---------------------------------------------------------------------------

def dmrg2(Ws, mps, bonds, chi_max, n_sweeps, tol=1e-10, krylov_max=KRYLOV_MAX_REF,
          cutoff=SVD_ABS_CUTOFF, rel_cutoff=0.0, record=None):
    """Two-site DMRG with the reference's sweep order and returned energy.

    Returns (energy, mps, bonds, sweep_energies).  ``record`` (dict) collects
    per-bond energies and Krylov dimensions.
    """
    L = len(Ws)
    D = Ws[0].shape[0]
    dt = np.result_type(mps[0], Ws[0])
    mps = [np.array(A, dtype=dt) for A in mps]
    bonds = list(bonds)
    Lenv = [None] * (L + 1)
    Renv = [None] * (L + 1)
    Lenv[0] = left_boundary(D, dt)
    Renv[L] = right_boundary(D, dt)
    for i in range(L - 1, -1, -1):                       # build_initial_environments
        Renv[i] = update_right_env(Renv[i + 1], mps[i], Ws[i])
    stats = {} if record is None else record
    bond_E = stats.setdefault('bond_energies', [])

    def optimize(site, direction):
        theta = np.einsum('asm,mtb->astb', mps[site], mps[site + 1])
        mv = lambda x: apply_heff2(Lenv[site], Renv[site + 2], Ws[site], Ws[site + 1], x)
        E, th = lanczos(mv, theta, site, krylov_max, stats)
        Al, Ar, S, disc = svd_split2(th, chi_max, direction, cutoff, rel_cutoff)
        mps[site], mps[site + 1] = Al, Ar
        bonds[site + 1] = Al.shape[2]
        stats['max_discarded'] = max(stats.get('max_discarded', 0.0), disc)
        bond_E.append(E)
        return E

    energy, e_prev = 0.0, 0.0
    sweep_E = []
    for sweep in range(n_sweeps):
        for site in range(L - 1):
            optimize(site, 'R')
            Lenv[site + 1] = update_left_env(Lenv[site], mps[site], Ws[site])
        for site in range(L - 2, -1, -1):
            energy = optimize(site, 'L')
            Renv[site + 1] = update_right_env(Renv[site + 2], mps[site + 1], Ws[site + 1])
        sweep_E.append(energy)
        if abs(energy - e_prev) < tol and sweep > 0:
            break
        e_prev = energy
    return energy, mps, bonds, sweep_E


# ---------------------------------------------------------------------------
# Single-site sweep (dmrg_gpu_impl.h:1118-1482)
# ---------------------------------------------------------------------------

def dmrg1(Ws, mps, bonds, chi_max, n_sweeps, tol=1e-10, krylov_max=KRYLOV_MAX_REF,
          cutoff=SVD_ABS_CUTOFF, rel_cutoff=0.0, record=None):
    """Single-site DMRG: SVD of (cL*d, cR) for 'R' with S Vh absorbed into the
    next site, SVD of (cL, d*cR) for 'L' with U S absorbed into the previous
    site; the end sites are optimised without an SVD."""
    L = len(Ws)
    D = Ws[0].shape[0]
    dt = np.result_type(mps[0], Ws[0])
    mps = [np.array(A, dtype=dt) for A in mps]
    bonds = list(bonds)
    Lenv = [None] * (L + 1)
    Renv = [None] * (L + 1)
    Lenv[0] = left_boundary(D, dt)
    Renv[L] = right_boundary(D, dt)
    for i in range(L - 1, -1, -1):
        Renv[i] = update_right_env(Renv[i + 1], mps[i], Ws[i])
    stats = {} if record is None else record
    bond_E = stats.setdefault('bond_energies', [])

    def solve(site):
        mv = lambda x: apply_heff1(Lenv[site], Renv[site + 1], Ws[site], x)
        E, th = lanczos(mv, mps[site], site, krylov_max, stats)
        bond_E.append(E)
        return E, th

    def split(site, th, direction):
        cL, d, cR = th.shape
        if direction == 'R':
            M = th.reshape(cL * d, cR, order='F')
        else:
            M = th.reshape(cL, d * cR, order='F')
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        k = min(len(S), chi_max)
        nk = truncation_rank(S, k, cutoff, rel_cutoff)
        U, S, Vh = U[:, :nk], S[:nk], Vh[:nk]
        if direction == 'R':
            mps[site] = U.reshape(cL, d, nk, order='F')
            SV = S[:, None] * Vh
            mps[site + 1] = np.tensordot(SV, mps[site + 1], axes=([1], [0]))
            bonds[site + 1] = nk
        else:
            mps[site] = Vh.reshape(nk, d, cR, order='F')
            US = U * S[None, :]
            mps[site - 1] = np.tensordot(mps[site - 1], US, axes=([2], [0]))
            bonds[site] = nk

    energy, e_prev = 0.0, 0.0
    sweep_E = []
    for sweep in range(n_sweeps):
        for site in range(L - 1):
            _, th = solve(site)
            split(site, th, 'R')
            Lenv[site + 1] = update_left_env(Lenv[site], mps[site], Ws[site])
        _, th = solve(L - 1)
        mps[L - 1] = th
        for site in range(L - 1, 0, -1):
            _, th = solve(site)
            split(site, th, 'L')
            Renv[site] = update_right_env(Renv[site + 1], mps[site], Ws[site])
        energy, th = solve(0)
        mps[0] = th
        sweep_E.append(energy)
        if abs(energy - e_prev) < tol and sweep > 0:
            break
        e_prev = energy
    return energy, mps, bonds, sweep_E
