"""Argument checks of the engine boundary that run before any device call, so
they hold on any machine: the reference's constructor takes the same
(L, d, chi_max, D) arguments (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:28-31)."""
import numpy as np
import pytest


def test_dtype_spellings(dmrg_amd):
    for f in ("f64", "float64", float, np.float64, dmrg_amd.DTYPE_F64):
        assert dmrg_amd.dtype_code(f) == dmrg_amd.DTYPE_F64
    for c in ("c128", "complex128", complex, np.complex128, dmrg_amd.DTYPE_C128):
        assert dmrg_amd.dtype_code(c) == dmrg_amd.DTYPE_C128
    for bad in ("f32", "bf16", np.float32, np.complex64, 2, True, None):
        with pytest.raises(ValueError):
            dmrg_amd.dtype_code(bad)


@pytest.mark.parametrize("args, kw, msg", [
    ((1, 2, 8, 5), {}, b"L must be"),
    ((4, 0, 8, 5), {}, b"physical dimension"),
    ((4, 2, 0, 5), {}, b"chi_max"),
    ((4, 2, 8, 0), {}, b"chi_max"),
    ((4, 2, 8, 5), {"variant": 3}, b"variant"),
    ((4, 2, 8, 5), {"krylov_max": 0}, b"krylov_max"),
    ((4, 2, 8, 5), {"krylov_max": 129}, b"krylov_max"),
    ((4, 2, 1025, 5), {}, b"chi_max * d"),
    ((4, 2, 8, 5), {"dtype": "c128", "precision": 1}, b"FP32 matvec"),
    ((4, 2, 8, 5), {"variant": 1, "precision": 2}, b"FP32 matvec"),
])
def test_create_rejects_bad_arguments(dmrg_amd, args, kw, msg):
    with pytest.raises(dmrg_amd.DMRGError) as ei:
        dmrg_amd.DMRG(*args, **kw)
    assert ei.value.code == dmrg_amd.ERR_BAD_ARG
    assert msg.decode() in str(ei.value)


def _unbound_engine(dmrg_amd, L=4, d=2, D=5):
    """A DMRG object with no library handle: only the binding's own checks,
    which run before any library call, can be exercised on it."""
    e = dmrg_amd.DMRG.__new__(dmrg_amd.DMRG)
    e.L, e.d, e.D, e.dtype, e._h = L, d, D, dmrg_amd.DTYPE_F64, None
    return e


def test_binding_checks_tensor_shapes(dmrg_amd):
    e = _unbound_engine(dmrg_amd)
    W = np.zeros((5, 2, 2, 5))
    with pytest.raises(ValueError, match="4 sites"):
        e.set_mpo([W] * 3)
    with pytest.raises(ValueError, match="site 2"):
        e.set_mpo([W, W, np.zeros((5, 3, 3, 5)), W])
    with pytest.raises(ValueError, match="site 1"):  # d = 1: the library would read past it
        e.set_mps([np.zeros((1, 2, 2)), np.zeros((2, 1, 2))])
    with pytest.raises(ValueError, match="site 0"):
        e.set_mps([np.zeros((2, 2))])
    with pytest.raises(ValueError, match="5 tensors"):
        e.set_mps([np.zeros((1, 2, 1))] * 5)


def _unbound_ops(dmrg_amd):
    o = dmrg_amd.Ops.__new__(dmrg_amd.Ops)
    o._h = None
    return o


def _arr(shape, dtype=np.float64):
    """Shape and dtype of a device array, without an allocation."""
    from types import SimpleNamespace
    return SimpleNamespace(shape=tuple(shape), dtype=np.dtype(dtype), ptr=None)


def test_op_wrappers_refuse_mismatched_device_arrays(dmrg_amd):
    """Each op reads and writes exactly the extents its sizes imply; arrays
    that disagree are refused before the library call."""
    o = _unbound_ops(dmrg_amd)
    cL, cR, d, D = 4, 3, 2, 5
    L, R, W = _arr((cL, D, cL)), _arr((cR, D, cR)), _arr((D, d, d, D))
    th, out = _arr((cL, d, d, cR)), _arr((cL, d, d, cR))
    cases = [
        (lambda: o.heff2(L, R, W, W, th, _arr((cL, d, d, cR + 1))), "out: shape"),
        (lambda: o.heff2(L, _arr((cR, D - 1, cR)), W, W, th, out), "R: shape"),
        (lambda: o.heff2(L, R, W, _arr((D, d, d, D), np.complex128), th, out), "W2: dtype"),
        (lambda: o.lanczos2(L, R, _arr((D, 3, 3, D)), W, th), "W1: shape"),
        (lambda: o.heff1(L, R, W, _arr((cL, d, cR)), _arr((cL, d, d, cR))), "out: shape"),
        (lambda: o.env_left(L, _arr((cL, d, 6)), W, _arr((6, D, 5))), "Lnew: shape"),
        (lambda: o.env_right(R, _arr((6, d, cR)), W, _arr((6, D, 6), np.complex128)), "Rnew: dtype"),
        (lambda: o.gemm(0, 0, _arr((7, 5)), _arr((4, 9)), _arr((7, 9))), "B: shape"),
        (lambda: o.gemm(1, 0, _arr((5, 7)), _arr((5, 9)), _arr((8, 9))), "A: shape"),
        (lambda: o.gemm_f32(0, 1, _arr((7, 5), np.float32), _arr((9, 5), np.float32),
                            _arr((7, 9))), "C: dtype"),
        (lambda: o.heig(_arr((6, 5))), "H: shape"),
        (lambda: o.svd_split(_arr((6,)), 2), "2-D"),
        (lambda: o.svd_split(_arr((6, 4), np.float32), 2), "f64 or c128"),
    ]
    for call, msg in cases:
        with pytest.raises(ValueError, match=msg):
            call()
