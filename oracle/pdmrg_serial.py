"""Restatement of the reference CPU path (cpu/pdmrg numpy primitives) --
oracle / CPU baseline, test-only.

This is the path that the reference's CPU implementation runs for the
two-site serial sweep (cpu/pdmrg/pdmrg/dmrg.py:114-178, identical to the
sweep body of cpu/pdmrg/pdmrg/tests/test_numerics.py:229-287), restated in
its own conventions:

* MPS tensor A[l, p, r] (C order), theta[l, p1, p2, r]
* environment L[bra, mpo, ket], R[bra, mpo, ket]
* MPO W[mpo_L, mpo_R, up(ket), down(bra)]

Functions follow cpu/pdmrg/pdmrg/numerics/effective_ham.py:21-117,
numerics/eigensolver.py:13-59, numerics/accurate_svd.py:82-110 and
environments/update.py:16-153.  scipy's ARPACK ``eigsh`` is the eigensolver,
as in the reference (scipy is a third-party dependency; uv.lock pins
scipy 1.13.1, this image has 1.15.3).

Pinning (tests/test_oracle_dmrg.py): apply_heff and update_left_env /
update_right_env against the H_eff and environment einsums of
test_numerics.py:112-126 / :162-192 written in these conventions
(test_apply_heff2_matches_reference_einsum, test_update_envs_match_reference_einsums),
and the serial sweep reaching the exact-diagonalisation energies of the
oracle's own DMRG (test_pdmrg_serial_matches_oracle; the reference test
asserts 1e-6 against quimb DMRG2, :229-287).
"""
import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh


def to_pdmrg_mpo(W):
    """Engine layout W[w, s, sp, wp] -> pdmrg W[w, wp, up=s, down=sp]."""
    return np.ascontiguousarray(np.transpose(W, (0, 3, 1, 2)))


def to_pdmrg_env(E):
    """Engine env E[ket, mpo, bra] -> pdmrg env [bra, mpo, ket]."""
    return np.ascontiguousarray(np.transpose(E, (2, 1, 0)))


def apply_heff(L_env, R_env, W_left, W_right, theta):
    """effective_ham.py:21-81: four tensordots."""
    X = np.tensordot(L_env, theta, axes=(2, 0))                      # a w s r e
    Y = np.tensordot(X, W_left, axes=([1, 2], [0, 2]))                # a r e m p
    Z = np.tensordot(Y, W_right, axes=([1, 3], [2, 0]))               # a e p n q
    return np.tensordot(Z, R_env, axes=([1, 3], [2, 1]))              # a p q f


def optimize_two_site(L_env, R_env, W_left, W_right, theta_init, max_iter=20, tol=1e-10):
    """eigensolver.py:13-59: eigsh(k=1, which='SA', v0=theta), with the
    reference's fallback of 10x maxiter and 10x tol on failure."""
    shape = theta_init.shape
    dim = int(np.prod(shape))
    dtype = np.result_type(theta_init.dtype, L_env.dtype, W_left.dtype)

    def mv(v):
        return apply_heff(L_env, R_env, W_left, W_right, v.reshape(shape)).ravel()

    H = LinearOperator((dim, dim), matvec=mv, dtype=dtype)
    v0 = theta_init.ravel().astype(dtype)
    try:
        E, V = eigsh(H, k=1, v0=v0, which='SA', maxiter=max_iter, tol=tol)
    except Exception:
        E, V = eigsh(H, k=1, v0=v0, which='SA', maxiter=max_iter * 10, tol=tol * 10)
    return float(E[0].real), V[:, 0].reshape(shape)


def truncated_svd(M, max_bond):
    """accurate_svd.py:82-110: keep min(k, max_bond), no cutoff."""
    U, S, Vh = np.linalg.svd(M, full_matrices=False)
    k = max(min(len(S), max_bond), 1)
    return U[:, :k], S[:k], Vh[:k, :], float(np.sum(S[k:] ** 2))


def update_left_env(L_env, A, W):
    """update.py:16-80."""
    X = np.tensordot(L_env, A, axes=(2, 0))
    Y = np.tensordot(X, W, axes=([1, 2], [0, 2]))
    L_new = np.tensordot(Y, A.conj(), axes=([0, 3], [0, 1]))
    return np.transpose(L_new, (2, 1, 0))


def update_right_env(R_env, B, W):
    """update.py:83-127."""
    X = np.tensordot(R_env, B, axes=(2, 2))
    Y = np.tensordot(X, W, axes=([1, 3], [1, 2]))
    R_new = np.tensordot(Y, B.conj(), axes=([0, 3], [2, 1]))
    return np.transpose(R_new, (2, 1, 0))


def boundary_envs(D, dtype):
    """Left boundary selects MPO row 0, right boundary column D-1 (engine
    MPO convention; update.py:130-153 uses index 0 / -1 the same way)."""
    L = np.zeros((1, D, 1), dtype)
    L[0, 0, 0] = 1.0
    R = np.zeros((1, D, 1), dtype)
    R[0, D - 1, 0] = 1.0
    return L, R


def bond_step(A, Ws, L_envs, R_envs, i, direction, max_bond, max_iter=20, tol=1e-10):
    """One two-site update of dmrg.py:114-178 at bond (i, i+1)."""
    theta = np.einsum('ijk,klm->ijlm', A[i], A[i + 1])
    E, th = optimize_two_site(L_envs[i], R_envs[i + 1], Ws[i], Ws[i + 1], theta,
                              max_iter=max_iter, tol=tol)
    cl, d1, d2, cr = th.shape
    U, S, Vh, _ = truncated_svd(th.reshape(cl * d1, d2 * cr), max_bond)
    if direction == 'right':
        A[i] = U.reshape(cl, d1, -1)
        A[i + 1] = (np.diag(S) @ Vh).reshape(-1, d2, cr)
        L_envs[i + 1] = update_left_env(L_envs[i], A[i], Ws[i])
    else:
        A[i + 1] = Vh.reshape(-1, d2, cr)
        A[i] = (U @ np.diag(S)).reshape(cl, d1, -1)
        R_envs[i] = update_right_env(R_envs[i + 1], A[i + 1], Ws[i + 1])
    return E


def serial_dmrg(A, Ws, max_bond, n_sweeps, max_iter=20, tol=1e-10):
    """Two-site serial DMRG with the reference CPU sweep (test_numerics.py:
    229-287).  A: list of (l, p, r) tensors; Ws: pdmrg-layout MPO.
    R_envs[i] holds the environment right of site i (sites i+1..L-1)."""
    L = len(A)
    A = [np.array(a) for a in A]
    D = Ws[0].shape[0]
    dt = np.result_type(A[0], Ws[0])
    Lb, Rb = boundary_envs(D, dt)
    L_envs = {0: Lb}
    R_envs = {L - 1: Rb}
    for i in range(L - 2, -1, -1):
        R_envs[i] = update_right_env(R_envs[i + 1], A[i + 1], Ws[i + 1])
    E = 0.0
    for _ in range(n_sweeps):
        for i in range(L - 1):
            E = bond_step(A, Ws, L_envs, R_envs, i, 'right', max_bond, max_iter, tol)
        for i in range(L - 2, -1, -1):
            E = bond_step(A, Ws, L_envs, R_envs, i, 'left', max_bond, max_iter, tol)
    return E, A
