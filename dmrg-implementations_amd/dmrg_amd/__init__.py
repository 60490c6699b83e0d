"""Python binding of the MI355X DMRG sweep engine (ctypes over libdmrg_amd.so).

Mirrors the reference's engine API (gpu-rocm/dmrg2-gpu/src/dmrg2_gpu.h:28-47):

    eng = DMRG(L, d, chi_max, D, tol=1e-10, dtype="f64", variant=2)
    eng.set_mpo(W)                 # list of (D, d, d, D) arrays, W[w, s, sp, wp]
    eng.initialize_mps_random(seed=1, scale=0.1) / initialize_mps_neel() / ...
    E = eng.run(n_sweeps)
    A = eng.get_mps()              # list of (cL, d, cR) arrays

Array conventions: logical numpy indices; the flat device layout is the
Fortran-order flattening, i.e. the reference's column-major layout.

The shared library is built by dmrg-implementations_amd/build.py.  Nothing
here falls back to a CPU implementation: if the library is missing, loading
raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmrg_amd.so")

OK, ERR_HIP, ERR_NO_CONV, ERR_BAD_ARG, ERR_NONFINITE, ERR_STATE, ERR_OOM = range(7)
DTYPE_F64, DTYPE_C128 = 0, 1
SINGLE_SITE, TWO_SITE = 1, 2
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2

EXPORTED = [
    "dmrg_default_options", "dmrg_create", "dmrg_destroy", "dmrg_set_mpo",
    "dmrg_init_mps_random", "dmrg_init_mps_product", "dmrg_init_mps_neel", "dmrg_set_mps",
    "dmrg_run", "dmrg_run_batch", "dmrg_run_pdmrg", "dmrg_get_energy", "dmrg_get_sweep_energies", "dmrg_get_bond_dims",
    "dmrg_get_mps", "dmrg_get_env", "dmrg_expectation", "dmrg_get_stats", "dmrg_reset_stats", "dmrg_last_error",
    "dmrg_ctx_create", "dmrg_ctx_destroy", "dmrg_op_heff2", "dmrg_op_heff2_f32", "dmrg_op_heff1", "dmrg_op_env_left",
    "dmrg_op_env_left_noalloc", "dmrg_op_env_right", "dmrg_op_svd_split", "dmrg_op_rsvd_split", "dmrg_op_heig", "dmrg_op_tridiag_lowest", "dmrg_op_lanczos2",
    "dmrg_op_gemm", "dmrg_op_gemm_f32", "dmrg_dev_malloc", "dmrg_dev_free", "dmrg_memcpy_h2d", "dmrg_memcpy_d2h", "dmrg_memcpy_d2d",
    "dmrg_dev_synchronize", "dmrg_device_count",
]


class DMRGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dmrg error %d: %s" % (code, msg))
        self.code = code


class Options(C.Structure):
    _fields_ = [("krylov_max", C.c_int), ("lanczos_tol", C.c_double), ("eig_tol", C.c_double),
                ("svd_cutoff", C.c_double), ("svd_rel_cutoff", C.c_double),
                ("precision", C.c_int), ("profile", C.c_int), ("device", C.c_int),
                ("svd_max_sweeps", C.c_int), ("svd_tol", C.c_double),
                ("svd_warm", C.c_int), ("fp32_eig_rel", C.c_double),
                ("svd_subspace", C.c_int), ("svd_tol_mixed", C.c_double),
                ("env_precision", C.c_int), ("svd_rsvd", C.c_int),
                ("rsvd_oversample", C.c_int), ("rsvd_power", C.c_int), ("graph", C.c_int),
                ("autotune", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("lanczos_ms", C.c_double), ("heff_ms", C.c_double), ("svd_ms", C.c_double),
                ("env_ms", C.c_double), ("sweep_s_last", C.c_double), ("heff_flop", C.c_double),
                ("env_flop", C.c_double), ("matvecs", C.c_longlong),
                ("lanczos_calls", C.c_longlong), ("svd_calls", C.c_longlong),
                ("svd_sweeps", C.c_longlong), ("sweeps_done", C.c_int),
                ("max_discarded", C.c_double), ("env_build_s", C.c_double),
                ("sweeps_s_total", C.c_double), ("svd_warm_starts", C.c_longlong),
                ("svd_bip_sweeps", C.c_longlong), ("svd_oa_steps", C.c_longlong),
                ("trunc_sum_last", C.c_double), ("rsvd_calls", C.c_longlong),
                ("graph_launches", C.c_longlong), ("graph_captures", C.c_longlong),
                ("tuned_shapes", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libdmrg_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libdmrg_amd.so not built: run dmrg-implementations_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, i, d, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
    sig = {
        "dmrg_default_options": (None, [C.POINTER(Options)]),
        "dmrg_create": (i, [i, i, i, i, d, i, i, C.POINTER(Options), C.POINTER(vp)]),
        "dmrg_destroy": (None, [vp]),
        "dmrg_set_mpo": (i, [vp, C.POINTER(vp)]),
        "dmrg_init_mps_random": (i, [vp, u64, d]),
        "dmrg_init_mps_product": (i, [vp]),
        "dmrg_init_mps_neel": (i, [vp]),
        "dmrg_set_mps": (i, [vp, i, vp, i, i]),
        "dmrg_run": (i, [vp, i, i, C.POINTER(d)]),
        "dmrg_run_batch": (i, [vp, i, i, i, vp, vp]),
        "dmrg_run_pdmrg": (i, [vp, i, i, i, i, i, i, C.POINTER(d)]),
        "dmrg_get_energy": (i, [vp, C.POINTER(d)]),
        "dmrg_get_sweep_energies": (i, [vp, C.POINTER(d), i, C.POINTER(i)]),
        "dmrg_get_bond_dims": (i, [vp, C.POINTER(i)]),
        "dmrg_get_mps": (i, [vp, i, vp]),
        "dmrg_get_env": (i, [vp, i, i, vp, C.POINTER(i)]),
        "dmrg_expectation": (i, [vp, C.POINTER(d), C.POINTER(d)]),
        "dmrg_get_stats": (i, [vp, C.POINTER(Stats)]),
        "dmrg_reset_stats": (i, [vp]),
        "dmrg_last_error": (C.c_char_p, []),
        "dmrg_ctx_create": (i, [i, C.POINTER(vp)]),
        "dmrg_ctx_destroy": (None, [vp]),
        "dmrg_op_heff2": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp, vp]),
        "dmrg_op_heff2_f32": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp, vp]),
        "dmrg_op_heff1": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp]),
        "dmrg_op_env_left": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp]),
        "dmrg_op_env_left_noalloc": (i, [vp, i, i, i, i, i, vp, vp, vp, vp]),
        "dmrg_op_env_right": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp]),
        "dmrg_op_svd_split": (i, [vp, vp, i, i, i, i, d, d, i, vp, vp, vp, vp, C.POINTER(i)]),
        "dmrg_op_rsvd_split": (i, [vp, vp, i, i, i, i, i, i, u64, d, d, i, vp, vp, vp,
                                   C.POINTER(i)]),
        "dmrg_op_heig": (i, [vp, vp, i, i, vp, vp, vp, C.POINTER(i)]),
        "dmrg_op_tridiag_lowest": (i, [vp, i, C.POINTER(d), C.POINTER(d), C.POINTER(d),
                                       C.POINTER(d)]),
        "dmrg_op_lanczos2": (i, [vp, vp, i, i, i, i, i, vp, vp, vp, vp, vp, i, i, C.POINTER(d),
                                 C.POINTER(i)]),
        "dmrg_op_gemm": (i, [vp, vp, i, i, i, i, i, i, vp, vp, i, vp, i, vp, vp, i]),
        "dmrg_op_gemm_f32": (i, [vp, vp, i, i, i, i, i, i, vp, i, vp, i, vp, i]),
        "dmrg_dev_malloc": (i, [C.POINTER(vp), C.c_size_t]),
        "dmrg_dev_free": (i, [vp]),
        "dmrg_memcpy_h2d": (i, [vp, vp, C.c_size_t]),
        "dmrg_memcpy_d2h": (i, [vp, vp, C.c_size_t]),
        "dmrg_memcpy_d2d": (i, [vp, vp, C.c_size_t]),
        "dmrg_dev_synchronize": (i, []),
        "dmrg_device_count": (i, [C.POINTER(i)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise DMRGError(rc, lib().dmrg_last_error().decode())


def default_options(**kw):
    o = Options()
    lib().dmrg_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError("unknown option %r" % k)
        setattr(o, k, v)
    return o


def device_count():
    n = C.c_int(0)
    _check(lib().dmrg_device_count(C.byref(n)))
    return n.value


def dtype_code(dtype):
    """DTYPE_F64 or DTYPE_C128 for "f64"/"float64"/float/np.float64 or
    "c128"/"complex128"/complex/np.complex128 (or the codes themselves);
    any other type raises ValueError rather than running in another one."""
    if isinstance(dtype, (int, np.integer)) and not isinstance(dtype, bool):
        if int(dtype) in (DTYPE_F64, DTYPE_C128):
            return int(dtype)
    elif isinstance(dtype, str):
        if dtype in ("f64", "float64"):
            return DTYPE_F64
        if dtype in ("c128", "complex128"):
            return DTYPE_C128
    elif dtype is not None and not isinstance(dtype, bool):  # np.dtype(None) is float64
        try:
            nd = np.dtype(dtype)
        except TypeError:
            nd = None
        if nd == np.float64:
            return DTYPE_F64
        if nd == np.complex128:
            return DTYPE_C128
    raise ValueError("dtype must be f64 or c128, got %r" % (dtype,))


def _np_dtype(dtype):
    return np.float64 if dtype == DTYPE_F64 else np.complex128


def _flat(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=_np_dtype(dtype)).ravel(order="F"))


class DMRG:
    """Handle on one engine (two-site or single-site, FP64 or complex128)."""

    def __init__(self, L, d, chi_max, D, tol=1e-10, dtype="f64", variant=TWO_SITE, **opts):
        self.dtype = dtype_code(dtype)
        self.L, self.d, self.chi_max, self.D = L, d, chi_max, D
        self.opts = default_options(**opts)
        h = C.c_void_p()
        _check(lib().dmrg_create(L, d, chi_max, D, tol, self.dtype, variant, C.byref(self.opts),
                                 C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().dmrg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state setup -------------------------------------------------------
    def set_mpo(self, Ws):
        if len(Ws) != self.L:
            raise ValueError("set_mpo: %d tensors for %d sites" % (len(Ws), self.L))
        for i, W in enumerate(Ws):
            if np.shape(W) != (self.D, self.d, self.d, self.D):
                raise ValueError("set_mpo: site %d has shape %s, expected %s"
                                 % (i, np.shape(W), (self.D, self.d, self.d, self.D)))
        flats = [_flat(W, self.dtype) for W in Ws]
        ptrs = (C.c_void_p * self.L)(*[f.ctypes.data for f in flats])
        _check(lib().dmrg_set_mpo(self._h, ptrs))

    def initialize_mps_random(self, seed=1, scale=0.1):
        _check(lib().dmrg_init_mps_random(self._h, C.c_uint64(seed), scale))

    def initialize_mps_product(self):
        _check(lib().dmrg_init_mps_product(self._h))

    def initialize_mps_neel(self):
        _check(lib().dmrg_init_mps_neel(self._h))

    def set_mps(self, mps):
        """Set sites 0 .. len(mps)-1; each A has shape (cL, d, cR)."""
        if len(mps) > self.L:
            raise ValueError("set_mps: %d tensors for %d sites" % (len(mps), self.L))
        for i, A in enumerate(mps):
            if np.ndim(A) != 3 or np.shape(A)[1] != self.d:
                raise ValueError("set_mps: site %d has shape %s, expected (cL, %d, cR)"
                                 % (i, np.shape(A), self.d))
        for i, A in enumerate(mps):
            f = _flat(A, self.dtype)
            _check(lib().dmrg_set_mps(self._h, i, f.ctypes.data, A.shape[0], A.shape[2]))

    # --- run / query -------------------------------------------------------
    def run(self, n_sweeps, quiet=True):
        e = C.c_double(0)
        _check(lib().dmrg_run(self._h, n_sweeps, 1 if quiet else 0, C.byref(e)))
        return e.value

    def run_pdmrg(self, segments=2, outer=10, local=2, warmup=1, polish=2, quiet=True):
        """Segment-parallel two-site sweep (dmrg_run_pdmrg)."""
        e = C.c_double(0)
        _check(lib().dmrg_run_pdmrg(self._h, segments, outer, local, warmup, polish,
                                    1 if quiet else 0, C.byref(e)))
        return e.value

    def get_energy(self):
        e = C.c_double(0)
        _check(lib().dmrg_get_energy(self._h, C.byref(e)))
        return e.value

    def sweep_energies(self):
        buf = (C.c_double * 1024)()
        n = C.c_int(0)
        _check(lib().dmrg_get_sweep_energies(self._h, buf, 1024, C.byref(n)))
        return list(buf[:min(n.value, 1024)])

    def bond_dims(self):
        buf = (C.c_int * (self.L + 1))()
        _check(lib().dmrg_get_bond_dims(self._h, buf))
        return list(buf)

    def chi_L(self, site):
        return self.bond_dims()[site]

    def chi_R(self, site):
        return self.bond_dims()[site + 1]

    def get_mps(self, site=None):
        b = self.bond_dims()
        sites = range(self.L) if site is None else [site]
        out = []
        for i in sites:
            a = np.empty(b[i] * self.d * b[i + 1], dtype=_np_dtype(self.dtype))
            _check(lib().dmrg_get_mps(self._h, i, a.ctypes.data))
            out.append(a.reshape(b[i], self.d, b[i + 1], order="F"))
        return out if site is None else out[0]

    def get_env(self, left, idx):
        """Environment (chi, D, chi) in the (ket, mpo, bra) layout."""
        c = C.c_int(0)
        _check(lib().dmrg_get_env(self._h, 1 if left else 0, idx, None, C.byref(c)))
        a = np.empty(c.value * self.D * c.value, dtype=_np_dtype(self.dtype))
        _check(lib().dmrg_get_env(self._h, 1 if left else 0, idx, a.ctypes.data, C.byref(c)))
        return a.reshape(c.value, self.D, c.value, order="F")

    def expectation(self):
        e, n = C.c_double(0), C.c_double(0)
        _check(lib().dmrg_expectation(self._h, C.byref(e), C.byref(n)))
        return e.value, n.value

    def stats(self):
        s = Stats()
        _check(lib().dmrg_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        _check(lib().dmrg_reset_stats(self._h))


# ---------------------------------------------------------------------------
# Device buffers and operator-level entry points
# ---------------------------------------------------------------------------

class DeviceArray:
    """A device allocation holding a logical numpy array (Fortran layout)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(lib().dmrg_dev_malloc(C.byref(p), max(self.nbytes, 8)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.asarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        flat = np.ascontiguousarray(a.ravel(order="F"))
        _check(lib().dmrg_memcpy_h2d(d.ptr, flat.ctypes.data, d.nbytes))
        return d

    def numpy(self):
        flat = np.empty(int(np.prod(self.shape)), dtype=self.dtype)
        _check(lib().dmrg_memcpy_d2h(flat.ctypes.data, self.ptr, self.nbytes))
        return flat.reshape(self.shape, order="F")

    def free(self):
        if self.ptr:
            lib().dmrg_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ops:
    """Operator-level calls on device buffers (one scratch context)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().dmrg_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().dmrg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _dt(arr):
        return dtype_code(arr.dtype)

    @staticmethod
    def _expect(dtype, **arrays):
        """Every named device array must have the given shape and dtype: the
        library reads and writes exactly those extents, so a mismatch is
        refused here rather than becoming an out-of-bounds device access."""
        for name, (arr, shape) in arrays.items():
            if tuple(arr.shape) != tuple(shape):
                raise ValueError("%s: shape %s, expected %s" % (name, tuple(arr.shape), tuple(shape)))
            if np.dtype(arr.dtype) != np.dtype(dtype):
                raise ValueError("%s: dtype %s, expected %s" % (name, arr.dtype, np.dtype(dtype)))

    def _two_site(self, L, R, W1, W2, theta, out=None):
        cL, D = L.shape[0], L.shape[1]
        cR, d = R.shape[0], theta.shape[1]
        arrays = dict(L=(L, (cL, D, cL)), R=(R, (cR, D, cR)), W1=(W1, (D, d, d, D)),
                      W2=(W2, (D, d, d, D)), theta=(theta, (cL, d, d, cR)))
        if out is not None:
            arrays["out"] = (out, (cL, d, d, cR))
        self._expect(theta.dtype, **arrays)
        return cL, cR, d, D

    def heff2(self, L, R, W1, W2, theta, out, stream=None):
        cL, cR, d, D = self._two_site(L, R, W1, W2, theta, out)
        _check(lib().dmrg_op_heff2(self._h, stream, self._dt(theta), cL, cR, d, D, L.ptr, R.ptr,
                                   W1.ptr, W2.ptr, theta.ptr, out.ptr))

    def heff2_f32(self, L, R, W1, W2, theta, out, prec=PREC_FP32, stream=None):
        """Mixed-mode two-site matvec: float32 L, R, theta, out; float64 W1, W2
        (fused on the host, rounded to FP32); FP32 MFMA or split-BF16 (prec)."""
        cL, D = L.shape[0], L.shape[1]
        cR, d = R.shape[0], theta.shape[1]
        self._expect(np.float32, L=(L, (cL, D, cL)), R=(R, (cR, D, cR)),
                     theta=(theta, (cL, d, d, cR)), out=(out, (cL, d, d, cR)))
        self._expect(np.float64, W1=(W1, (D, d, d, D)), W2=(W2, (D, d, d, D)))
        _check(lib().dmrg_op_heff2_f32(self._h, stream, prec, cL, cR, d, D, L.ptr, R.ptr, W1.ptr,
                                       W2.ptr, theta.ptr, out.ptr))

    def heff1(self, L, R, W, theta, out, stream=None):
        cL, D = L.shape[0], L.shape[1]
        cR, d = R.shape[0], theta.shape[1]
        self._expect(theta.dtype, L=(L, (cL, D, cL)), R=(R, (cR, D, cR)), W=(W, (D, d, d, D)),
                     theta=(theta, (cL, d, cR)), out=(out, (cL, d, cR)))
        _check(lib().dmrg_op_heff1(self._h, stream, self._dt(theta), cL, cR, d, D, L.ptr, R.ptr,
                                   W.ptr, theta.ptr, out.ptr))

    def env_left(self, L, A, W, Lnew, stream=None):
        ci, d, co = A.shape
        D = W.shape[0]
        self._expect(A.dtype, L=(L, (ci, D, ci)), A=(A, (ci, d, co)), W=(W, (D, d, d, D)),
                     Lnew=(Lnew, (co, D, co)))
        _check(lib().dmrg_op_env_left(self._h, stream, self._dt(A), ci, co, d, D, L.ptr, A.ptr,
                                      W.ptr, Lnew.ptr))

    def env_right(self, R, A, W, Rnew, stream=None):
        co, d, ci = A.shape
        D = W.shape[0]
        self._expect(A.dtype, R=(R, (ci, D, ci)), A=(A, (co, d, ci)), W=(W, (D, d, d, D)),
                     Rnew=(Rnew, (co, D, co)))
        _check(lib().dmrg_op_env_right(self._h, stream, self._dt(A), ci, co, d, D, R.ptr, A.ptr,
                                       W.ptr, Rnew.ptr))

    def svd_split(self, M, k, cutoff=1e-14, rel_cutoff=0.0, scale_left=False):
        """Truncated split of a device matrix M (m x n); returns numpy
        (left, right, sigma, nk)."""
        if len(M.shape) != 2:
            raise ValueError("svd_split: M must be 2-D, got shape %s" % (M.shape,))
        m, n = M.shape
        dt = M.dtype
        dtype_code(dt)
        left = DeviceArray((m, k), dt)
        right = DeviceArray((k, n), dt)
        sig = DeviceArray((k,), np.float64)
        nk = C.c_int(0)
        _check(lib().dmrg_op_svd_split(self._h, None, self._dt(M), m, n, k, cutoff, rel_cutoff,
                                       1 if scale_left else 0, M.ptr, left.ptr, right.ptr, sig.ptr,
                                       C.byref(nk)))
        nk = nk.value
        lf = left.numpy().ravel(order="F")[: m * nk].reshape(m, nk, order="F")
        rf = right.numpy().ravel(order="F")[: nk * n].reshape(nk, n, order="F")
        return lf, rf, sig.numpy()[:nk], nk

    def rsvd_split(self, M, k, oversample=10, power=0, seed=0x5EED0000, cutoff=1e-14,
                   rel_cutoff=0.0, scale_left=False):
        """Randomised truncated split of a device matrix M (m x n); returns
        numpy (left, right, nk)."""
        if len(M.shape) != 2:
            raise ValueError("rsvd_split: M must be 2-D, got shape %s" % (M.shape,))
        m, n = M.shape
        dt = M.dtype
        dtype_code(dt)
        left = DeviceArray((m, k), dt)
        right = DeviceArray((k, n), dt)
        nk = C.c_int(0)
        _check(lib().dmrg_op_rsvd_split(self._h, None, self._dt(M), m, n, k, oversample, power,
                                        seed, cutoff, rel_cutoff, 1 if scale_left else 0, M.ptr,
                                        left.ptr, right.ptr, C.byref(nk)))
        nk = nk.value
        lf = left.numpy().ravel(order="F")[: m * nk].reshape(m, nk, order="F")
        rf = right.numpy().ravel(order="F")[: nk * n].reshape(nk, n, order="F")
        return lf, rf, nk

    def heig(self, H):
        """Hermitian eigen-decomposition of a device matrix H (N x N); returns
        numpy (evals descending, evecs, sweeps)."""
        N = H.shape[0]
        self._expect(H.dtype, H=(H, (N, N)))
        dtype_code(H.dtype)
        ev = DeviceArray((N,), np.float64)
        V = DeviceArray((N, N), H.dtype)
        sw = C.c_int(0)
        _check(lib().dmrg_op_heig(self._h, None, self._dt(H), N, H.ptr, ev.ptr, V.ptr,
                                  C.byref(sw)))
        return ev.numpy(), V.numpy(), sw.value

    def gemm(self, opA, opB, A, B, Cm, alpha=1.0, beta=0.0, stream=None):
        """Cm = alpha op(A) op(B) + beta Cm on 2-D device arrays."""
        dt = self._dt(A)
        npd = _np_dtype(dt)
        M, N = Cm.shape
        K = A.shape[0] if opA in (1, 3) else A.shape[1]
        self._expect(npd, A=(A, (K, M) if opA in (1, 3) else (M, K)),
                     B=(B, (N, K) if opB in (1, 3) else (K, N)), C=(Cm, (M, N)))
        a = np.array([alpha], dtype=npd)
        b = np.array([beta], dtype=npd)
        _check(lib().dmrg_op_gemm(self._h, stream, dt, opA, opB, M, N, K, a.ctypes.data, A.ptr,
                                  A.shape[0], B.ptr, B.shape[0], b.ctypes.data, Cm.ptr, M))

    def gemm_f32(self, opA, opB, A, B, Cm, prec=PREC_FP32, stream=None):
        """FP32 Cm = op(A) op(B) (op 0 or 1) on float32 device arrays, FP32 MFMA
        or split-BF16 (prec)."""
        M, N = Cm.shape
        K = A.shape[0] if opA == 1 else A.shape[1]
        self._expect(np.float32, A=(A, (K, M) if opA == 1 else (M, K)),
                     B=(B, (N, K) if opB == 1 else (K, N)), C=(Cm, (M, N)))
        _check(lib().dmrg_op_gemm_f32(self._h, stream, prec, opA, opB, M, N, K, A.ptr, A.shape[0],
                                      B.ptr, B.shape[0], Cm.ptr, M))

    def tridiag_lowest(self, alpha, beta):
        k = len(alpha)
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        b = np.ascontiguousarray(beta if k > 1 else [0.0], dtype=np.float64)
        lam = C.c_double(0)
        y = np.zeros(k)
        dp = C.POINTER(C.c_double)
        _check(lib().dmrg_op_tridiag_lowest(self._h, k, a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                            C.byref(lam), y.ctypes.data_as(dp)))
        return lam.value, y

    def lanczos2(self, L, R, W1, W2, theta, krylov_max=100, site=0, stream=None):
        cL, cR, d, D = self._two_site(L, R, W1, W2, theta)
        e = C.c_double(0)
        nit = C.c_int(0)
        _check(lib().dmrg_op_lanczos2(self._h, stream, self._dt(theta), cL, cR, d, D, L.ptr, R.ptr,
                                      W1.ptr, W2.ptr, theta.ptr, krylov_max, site, C.byref(e),
                                      C.byref(nit)))
        return e.value, nit.value


def run_batch(engines, n_sweeps, threads=8):
    """Run independent DMRG engines concurrently (dmrg_run_batch); returns
    (energies, seconds) as numpy arrays in the order of `engines`."""
    import numpy as _np
    n = len(engines)
    hs = (C.c_void_p * max(n, 1))(*[e._h.value for e in engines])
    E = _np.zeros(max(n, 1))
    T = _np.zeros(max(n, 1))
    _check(lib().dmrg_run_batch(hs, n, n_sweeps, threads, E.ctypes.data, T.ctypes.data))
    return E[:n], T[:n]


# bytes of the library's fixed device scratch behind env_left_noalloc
NOALLOC_SCRATCH_BYTES = 16 << 20


def env_left_noalloc(L, A, W, Lnew, stream=None):
    """dmrg_op_env_left_noalloc: the left-environment update of
    Ops.env_left on the library's fixed device scratch (no allocation, no
    context), on the current device.  Calls that overlap on one GPU share
    that scratch and race."""
    ci, d, co = A.shape
    D = W.shape[0]
    Ops._expect(A.dtype, L=(L, (ci, D, ci)), A=(A, (ci, d, co)), W=(W, (D, d, d, D)),
                Lnew=(Lnew, (co, D, co)))
    _check(lib().dmrg_op_env_left_noalloc(stream, dtype_code(A.dtype), ci, co, d, D, L.ptr, A.ptr,
                                          W.ptr, Lnew.ptr))


def copy_device(dst, src):
    """Device-to-device copy of src into dst (same byte size)."""
    if dst.nbytes != src.nbytes:
        raise ValueError("copy_device: size mismatch")
    _check(lib().dmrg_memcpy_d2d(dst.ptr, src.ptr, src.nbytes))


def synchronize():
    _check(lib().dmrg_dev_synchronize())
