"""MPO builders and exact-diagonalisation Hamiltonians (oracle, test-only).

Layout
------
Every MPO tensor is returned as a numpy array ``W`` of logical shape
``(D, d, d, D)`` indexed ``W[w, s, sp, wp]`` where ``s`` is the ket (input)
physical index, ``sp`` the bra (output) index, ``w`` the left and ``wp`` the
right MPO bond.  Flattened in Fortran order this is exactly the reference's
device layout ``W[w + s*D + sp*D*d + wp*D*d*d]``
(gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:22-34).  ``W[w, s, sp, wp]`` holds
the matrix element ``<sp| O_{w,wp} |s>``.  The left boundary selects row
``w = 0`` and the right boundary column ``wp = D-1``
(gpu-rocm/dmrg2-gpu/src/dmrg2_gpu_impl.h:944-956).

Pinning
-------
* ``heisenberg_mpo`` / ``josephson_mpo`` / ``tfim_mpo`` /
  ``random_field_heisenberg_mpo``: contracted to a dense matrix by
  ``mpo_to_dense`` and compared element-wise with the Kronecker-product
  Hamiltonians below (built from the operator definitions, not from the MPO),
  and the ground energies compared with the reference's exact-diagonalisation
  known answers (test_dmrg2_gpu.cpp:208-211, :292-297).
* ``dense_*``: pinned by those same known answers (Heisenberg L=4/8,
  Josephson n_max=1 L=4/6) and by closed forms (two-site singlet -3/4, the
  all-up eigenstate (L-1)/4; the transverse-field Ising chain by its
  free-fermion ground energy, minus the singular-value sum of the bidiagonal
  matrix with h on the diagonal and J above it).
* ``spin_chain_mpo`` / ``ladder_mpo`` and ``dense_spin_chain`` /
  ``dense_ladder``: MPO against Kronecker form element-wise; ground energies
  against the reference's known answers (J1-J2 at the Majumdar-Ghosh point,
  -3L/8: test_dmrg2_gpu.cpp:344-346; J1-J2-J3: :390-394; ladder: :442-447),
  the Majumdar-Ghosh dimer closed form and J2 = J3 = 0 reducing to
  ``dense_heisenberg``.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# Spin-1/2 operators as <row=bra| op |col=ket>; basis 0 = up (Sz=+1/2).
_SP = np.array([[0.0, 1.0], [0.0, 0.0]])
_SM = np.array([[0.0, 0.0], [1.0, 0.0]])
_SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
_ID2 = np.eye(2)


def _set_block(W, w, wp, op):
    """W[w, s, sp, wp] = op[sp, s]  (op given as <bra|op|ket> matrix)."""
    W[w, :, :, wp] = np.asarray(op).T


def heisenberg_mpo(L, J=1.0):
    """Spin-1/2 XXX chain, open BC, D=5 (test_dmrg2_gpu.cpp:20-63)."""
    return random_field_heisenberg_mpo(L, np.zeros(L), J=J)


def random_field_heisenberg_mpo(L, h, J=1.0):
    """H = J sum_i S_i.S_{i+1} + sum_i h_i Sz_i, D=5, upper triangular.

    Same bond structure as the reference Heisenberg MPO
    (test_dmrg2_gpu.cpp:20-63); the on-site field sits in block (0, D-1).
    """
    h = np.asarray(h, dtype=float)
    assert h.shape == (L,)
    D = 5
    Ws = []
    for i in range(L):
        W = np.zeros((D, 2, 2, D))
        first, last = (i == 0), (i == L - 1)
        if not last:
            _set_block(W, 0, 0, _ID2)
            _set_block(W, 0, 1, J * _SP)
            _set_block(W, 0, 2, J * _SM)
            _set_block(W, 0, 3, J * _SZ)
        if not first:
            _set_block(W, 1, 4, 0.5 * _SM)
            _set_block(W, 2, 4, 0.5 * _SP)
            _set_block(W, 3, 4, _SZ)
            _set_block(W, 4, 4, _ID2)
        if h[i] != 0.0:
            _set_block(W, 0, 4, h[i] * _SZ)
        if first and last:  # L == 1: only the on-site term survives
            pass
        Ws.append(W)
    return Ws


def spin_chain_mpo(L, Js):
    """H = sum_{r=1..R} J_r sum_i S_i.S_{i+r}, open chain, R = len(Js).

    Finite-automaton MPO: state 0 = nothing placed; states (r, a, j) for the
    range r, the opening operator a in (S+, S-, Sz) and j = 0..r-1 identities
    placed since; state D-1 = complete.  D = 2 + 3R(R+1)/2 (11 for J1-J2,
    20 for J1-J2-J3; gpu-rocm/common/challenge_mpos.h:18-51,
    test_dmrg2_gpu.cpp:329-418 use the same bond dimensions)."""
    R = len(Js)
    D = 2 + 3 * R * (R + 1) // 2
    opens = (_SP, _SM, _SZ)
    closes = (0.5 * _SM, 0.5 * _SP, _SZ)
    Ws = []
    for _ in range(L):
        W = np.zeros((D, 2, 2, D))
        _set_block(W, 0, 0, _ID2)
        _set_block(W, D - 1, D - 1, _ID2)
        base = 1
        for r in range(1, R + 1):
            for a in range(3):
                s0 = base + a * r
                _set_block(W, 0, s0, opens[a])
                for j in range(r - 1):
                    _set_block(W, s0 + j, s0 + j + 1, _ID2)
                _set_block(W, s0 + r - 1, D - 1, Js[r - 1] * closes[a])
            base += 3 * r
        Ws.append(W)
    return Ws


def ladder_mpo(L, J_leg=1.0, J_rung=1.0):
    """Two-leg Heisenberg ladder, one rung per site (d = 4, s = 2 s1 + s2),
    H = J_leg sum_i (S1_i.S1_{i+1} + S2_i.S2_{i+1}) + J_rung sum_i S1_i.S2_i,
    D = 8: 0 start, 1-3 / 4-6 leg 1 / leg 2 opened by S+, S-, Sz, 7 done;
    the rung term on-site in block (0, 7) (gpu-rocm/common/challenge_mpos.h:126-170)."""
    I2 = _ID2
    S1 = [np.kron(o, I2) for o in (_SP, _SM, _SZ)]
    S2 = [np.kron(I2, o) for o in (_SP, _SM, _SZ)]
    rung = J_rung * (0.5 * (np.kron(_SP, _SM) + np.kron(_SM, _SP)) + np.kron(_SZ, _SZ))
    partner, scale = (1, 0, 2), (0.5, 0.5, 1.0)
    D = 8
    Ws = []
    for _ in range(L):
        W = np.zeros((D, 4, 4, D))
        _set_block(W, 0, 0, np.eye(4))
        _set_block(W, D - 1, D - 1, np.eye(4))
        _set_block(W, 0, D - 1, rung)
        for leg, S in enumerate((S1, S2)):
            for a in range(3):
                st = 1 + 3 * leg + a
                _set_block(W, 0, st, S[a])
                _set_block(W, st, D - 1, J_leg * scale[a] * S[partner[a]])
        Ws.append(W)
    return Ws


def tfim_mpo(L, J=1.0, hx=1.0):
    """H = -J sum sz_i sz_{i+1} - hx sum sx_i (Pauli), D=3
    (test_dmrg2_gpu.cpp:76-120)."""
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    sz = np.array([[1.0, 0.0], [0.0, -1.0]])
    D = 3
    Ws = []
    for i in range(L):
        W = np.zeros((D, 2, 2, D))
        if i != L - 1:
            _set_block(W, 0, 0, _ID2)
            _set_block(W, 0, 1, -J * sz)
        _set_block(W, 0, 2, -hx * sx)
        if i != 0:
            _set_block(W, 1, 2, sz)
            _set_block(W, 2, 2, _ID2)
        Ws.append(W)
    return Ws


def josephson_ops(n_max, E_C, n_g=0.0):
    d = 2 * n_max + 1
    q = np.arange(d) - n_max
    onsite = np.diag(E_C * q * q - n_g * q).astype(complex)
    up = np.zeros((d, d), complex)     # e^{i phi}: |n> -> |n+1>
    for k in range(d - 1):
        up[k + 1, k] = 1.0
    return d, onsite, up, up.conj().T


def josephson_mpo(L, n_max=2, E_J=2.0, E_C=1.0, phi_ext=np.pi / 3, n_g=0.0):
    """Josephson-junction array, complex, D=4 (test_dmrg2_gpu.cpp:122-191).

    H = sum_i (E_C n_i^2 - n_g n_i) - E_J sum_i cos(phi_i - phi_{i+1} + Phi),
    written with alpha = -E_J/2 e^{i Phi} on e^{i phi_i} e^{-i phi_{i+1}}.
    """
    d, onsite, up, down = josephson_ops(n_max, E_C, n_g)
    alpha = -0.5 * E_J * np.exp(1j * phi_ext)
    eye = np.eye(d, dtype=complex)
    D = 4
    Ws = []
    for i in range(L):
        W = np.zeros((D, d, d, D), complex)
        if i != L - 1:
            _set_block(W, 0, 0, eye)
            _set_block(W, 0, 1, alpha * up)
            _set_block(W, 0, 2, np.conj(alpha) * down)
        _set_block(W, 0, 3, onsite)
        if i != 0:
            _set_block(W, 1, 3, down)
            _set_block(W, 2, 3, up)
            _set_block(W, 3, 3, eye)
        Ws.append(W)
    return Ws


# ---------------------------------------------------------------------------
# Dense / sparse Hamiltonians for exact diagonalisation
# ---------------------------------------------------------------------------

def mpo_to_dense(Ws):
    """Contract an MPO (boundary rows w=0, wp=D-1) to the dense matrix
    H[(sp_0..sp_{L-1}), (s_0..s_{L-1})], site 0 most significant."""
    D = Ws[0].shape[0]
    # T[w, bra, ket]
    T = np.zeros((D, 1, 1), dtype=np.result_type(*Ws))
    T[0, 0, 0] = 1.0
    for W in Ws:
        # new[wp, bra*d + sp, ket*d + s] = sum_w T[w, bra, ket] W[w, s, sp, wp]
        d = W.shape[1]
        new = np.einsum('wbk,wsxv->vbxks', T, W)
        nb, nk = T.shape[1] * d, T.shape[2] * d
        T = new.reshape(W.shape[3], nb, nk)
    return T[D - 1]


def _kron_chain(ops):
    out = sp.csr_matrix(np.ones((1, 1)))
    for o in ops:
        out = sp.kron(out, sp.csr_matrix(o), format='csr')
    return out


def _site_op(L, i, op, d):
    ops = [np.eye(d)] * L
    ops = list(ops)
    ops[i] = op
    return _kron_chain(ops)


def dense_heisenberg(L, J=1.0, h=None):
    """Sparse H = J sum S_i.S_{i+1} + sum h_i Sz_i from Kronecker products."""
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L - 1):
        for a, b, c in ((_SP, _SM, 0.5), (_SM, _SP, 0.5), (_SZ, _SZ, 1.0)):
            ops = [np.eye(2)] * L
            ops = list(ops)
            ops[i], ops[i + 1] = a, b
            H = H + (J * c) * _kron_chain(ops)
    if h is not None:
        for i in range(L):
            if h[i] != 0.0:
                H = H + h[i] * _site_op(L, i, _SZ, 2)
    return H


def _dense_ss(n, pairs):
    """sum over (i, j, c) in pairs of c S_i.S_j on n spins (Kronecker products)."""
    H = sp.csr_matrix((2 ** n, 2 ** n))
    for i, j, c in pairs:
        for a, b, f in ((_SP, _SM, 0.5), (_SM, _SP, 0.5), (_SZ, _SZ, 1.0)):
            ops = [np.eye(2)] * n
            ops = list(ops)
            ops[i], ops[j] = a, b
            H = H + (c * f) * _kron_chain(ops)
    return H


def dense_spin_chain(L, Js):
    """H = sum_r J_r sum_i S_i.S_{i+r} from the operator definitions."""
    return _dense_ss(L, [(i, i + r, J) for r, J in enumerate(Js, start=1) for i in range(L - r)])


def dense_ladder(L, J_leg=1.0, J_rung=1.0):
    """2L spins, rung i = spins (2i, 2i+1) (leg 1 first, matching s = 2 s1 + s2)."""
    pairs = [(2 * i, 2 * i + 1, J_rung) for i in range(L)]
    pairs += [(2 * i + leg, 2 * i + 2 + leg, J_leg) for i in range(L - 1) for leg in (0, 1)]
    return _dense_ss(2 * L, pairs)


def dense_tfim(L, J=1.0, hx=1.0):
    sx = np.array([[0.0, 1.0], [1.0, 0.0]])
    sz = np.array([[1.0, 0.0], [0.0, -1.0]])
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L - 1):
        ops = [np.eye(2)] * L
        ops = list(ops)
        ops[i], ops[i + 1] = sz, sz
        H = H - J * _kron_chain(ops)
    for i in range(L):
        H = H - hx * _site_op(L, i, sx, 2)
    return H


def dense_josephson(L, n_max=2, E_J=2.0, E_C=1.0, phi_ext=np.pi / 3, n_g=0.0):
    d, onsite, up, down = josephson_ops(n_max, E_C, n_g)
    H = sp.csr_matrix((d ** L, d ** L), dtype=complex)
    for i in range(L):
        H = H + _site_op(L, i, onsite, d)
    for i in range(L - 1):
        ops = [np.eye(d)] * L
        ops = list(ops)
        ops[i], ops[i + 1] = up, down
        hop = _kron_chain(ops)          # e^{i(phi_i - phi_{i+1})}
        H = H - 0.5 * E_J * (np.exp(1j * phi_ext) * hop
                             + np.exp(-1j * phi_ext) * hop.conj().T)
    return H


def ground_energy(H, dense_limit=4096):
    """Lowest eigenvalue of a Hermitian (sparse) matrix."""
    n = H.shape[0]
    if n <= dense_limit:
        A = H.toarray() if sp.issparse(H) else H
        return float(np.linalg.eigvalsh(A)[0])
    e = spla.eigsh(H, k=1, which='SA', tol=1e-14, maxiter=100000,
                   return_eigenvectors=False)
    return float(e[0].real)
