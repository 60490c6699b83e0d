"""The allocation-free left-environment update against dmrg_op_env_left.

dmrg_op_env_left_noalloc runs the kernels of dmrg_op_env_left with its two
intermediates in the library's fixed device scratch.  Two calls in a row, the
second on a smaller input (so the scratch holds the first call's larger
intermediates when the second starts), must each equal dmrg_op_env_left on the
same inputs bit for bit (same kernels, same order of operations), and equal
the oracle's einsum-pinned contraction to 1e-12.  Sizes past the scratch are
refused with BAD_ARG before any launch.
"""
import numpy as np
import pytest

from conftest import rand, rel_err
from oracle import dmrg as od

pytestmark = pytest.mark.gpu


def _case(dmrg_amd, rng, ci, co, d, D, cplx):
    W = rand(rng, (D, d, d, D), cplx)
    Lg = rand(rng, (ci, D, ci), cplx)
    A = rand(rng, (ci, d, co), cplx)
    dev = dmrg_amd.DeviceArray.from_numpy
    return (Lg, A, W), (dev(Lg), dev(A), dev(W))


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_env_left_noalloc_matches_env_left(dmrg_amd, ops, cplx):
    rng = np.random.default_rng(7 + cplx)
    # first call: FP64 256 x 256 (10 MiB of intermediates; c128 uses 160 x 192)
    shapes = [(160, 192, 2, 5), (48, 40, 2, 5)] if cplx else [(256, 256, 2, 5), (64, 48, 2, 5)]
    for ci, co, d, D in shapes:
        host, (Lg, A, W) = _case(dmrg_amd, rng, ci, co, d, D, cplx)
        dt = A.dtype
        ref = dmrg_amd.DeviceArray((co, D, co), dt)
        ops.env_left(Lg, A, W, ref)
        got = dmrg_amd.DeviceArray((co, D, co), dt)
        dmrg_amd.env_left_noalloc(Lg, A, W, got)
        dmrg_amd.synchronize()
        g, r = got.numpy(), ref.numpy()
        assert np.array_equal(g, r), (ci, co, np.max(np.abs(g - r)))
        assert rel_err(g, od.update_left_env(*host)) < 1e-12


def test_env_left_noalloc_refuses_oversize(dmrg_amd):
    """2 D d ci co scalars past the 16 MiB scratch: BAD_ARG, nothing launched."""
    ci = co = 512  # FP64 d=2 D=5: 2 x 20 MiB
    rng = np.random.default_rng(3)
    _, (Lg, A, W) = _case(dmrg_amd, rng, ci, co, 2, 5, False)
    out = dmrg_amd.DeviceArray((co, 5, co), np.float64)
    with pytest.raises(dmrg_amd.DMRGError) as ei:
        dmrg_amd.env_left_noalloc(Lg, A, W, out)
    assert "scratch" in str(ei.value)
