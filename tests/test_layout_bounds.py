"""CPU checks of the two-level GEMM views used by the contractions.

The views below restate the (n1, s1, s2) triples that contract.hip passes
to gemm_tensor; for every shape the largest address reached must be the last
element of the buffer (a bijection onto [0, size)), so no launch can read or
write outside its tensor.
"""
import itertools

import numpy as np
import pytest


def addr(dim, i):
    n1, s1, s2 = dim
    return (i % n1) * s1 + (i // n1) * s2


def all_addrs(view, R, Cn):
    r = np.arange(R)[:, None]
    c = np.arange(Cn)[None, :]
    return addr(view[0], r) + addr(view[1], c)


def views(kind, a, b, d, D):
    """(M, N, K, viewX, viewY, size) for each step-2 contraction."""
    if kind == "heff2":
        cL, cR, dd = a, b, d * d
        vx = ((cL, D, D * cL * dd), (D, 1, D * cL))
        vy = ((cL, 1, cL * dd), (dd, cL, cL * dd * cR))
        return cL * cR, dd * D, D * dd, vx, vy, D * dd * cL * cR
    if kind in ("env_left", "heff1"):
        ci, co = a, b
        vx = ((ci, D, D * ci * d), (D, 1, D * ci))
        vy = ((ci, 1, ci * d), (d, ci, ci * d * co))
        return ci * co, d * D, D * d, vx, vy, D * d * ci * co
    if kind == "env_right":
        ci, co = a, b
        vx = ((co, 1, co * d * D), (d, co, co * d))
        vy = ((co, 1, co * D * d), (D, co, co * D))
        return co * ci, D * d, d * D, vx, vy, D * d * ci * co
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["heff2", "heff1", "env_left", "env_right"])
@pytest.mark.parametrize("a,b,d,D", list(itertools.product([1, 3, 16], [1, 5, 16], [2, 5], [3, 5])))
def test_views_are_bijections(kind, a, b, d, D):
    M, N, K, vx, vy, size = views(kind, a, b, d, D)
    ax = all_addrs(vx, M, K).ravel()
    ay = all_addrs(vy, M, N).ravel()
    for arr in (ax, ay):
        assert arr.min() == 0 and arr.max() == size - 1
        assert np.unique(arr).size == size


@pytest.mark.parametrize("cR", [1, 3, 16, 33])
@pytest.mark.parametrize("d", [2, 3, 5])
def test_warm_basis_right_view_is_a_bijection(cR, d):
    """warm.hip warm_basis_right: the stored N x N eigenbasis X (N = d*cR) is
    read and written as a (d*N) x cR operand, row i = s + d*j -> s + N*j,
    column k -> d*k (the same view for the output).  Every address of the
    N x N buffer is reached exactly once."""
    N = d * cR
    v = ((d, 1, N), (1 << 30, d, 0))
    a = all_addrs(v, d * N, cR).ravel()
    assert a.min() == 0 and a.max() == N * N - 1
    assert np.unique(a).size == N * N == a.size
