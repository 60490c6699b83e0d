"""The C-ABI library loads on any machine and exports every symbol that
include/dmrg_amd.h declares (no compute call is made here)."""
import os
import re

from conftest import ROOT


def test_header_symbols_exported(dmrg_amd):
    hdr = open(os.path.join(ROOT, "include", "dmrg_amd.h")).read()
    declared = set(re.findall(r"\b(dmrg_[a-z0-9_]+)\s*\(", hdr))
    lib = dmrg_amd.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(dmrg_amd.EXPORTED) <= declared


def test_errors_without_device_are_reported(dmrg_amd):
    if dmrg_amd.device_count() > 0:
        return
    try:
        dmrg_amd.DMRG(4, 2, 8, 5)
    except dmrg_amd.DMRGError as e:
        assert e.code == dmrg_amd.ERR_HIP
    else:
        raise AssertionError("engine created without a device")


def test_null_outputs_rejected_before_any_device_call(dmrg_amd):
    """Output pointers that an entry point would write through are checked on
    the host first; the handle below is only a non-NULL stand-in, never
    dereferenced on these paths."""
    import ctypes as C

    lib = dmrg_amd.lib()
    stand_in = C.create_string_buffer(64)
    h = C.cast(stand_in, C.c_void_p)
    n = C.c_int(-1)
    assert lib.dmrg_get_sweep_energies(h, None, 4, C.byref(n)) == dmrg_amd.ERR_BAD_ARG
    assert n.value == -1
    nk = C.c_int(-1)
    data = C.create_string_buffer(64)
    buf = C.cast(data, C.c_void_p)
    for left, right in ((None, buf), (buf, None)):
        rc = lib.dmrg_op_svd_split(h, None, dmrg_amd.DTYPE_F64, 4, 4, 2, 1e-14, 0.0, 0, buf,
                                   left, right, None, C.byref(nk))
        assert rc == dmrg_amd.ERR_BAD_ARG
        assert b"NULL" in lib.dmrg_last_error()
    assert nk.value == -1


def test_batch_refuses_an_engine_listed_twice(dmrg_amd):
    """Both entries would run on two threads at once; the check precedes any
    use of the handles (stand-ins here, never dereferenced)."""
    import ctypes as C

    lib = dmrg_amd.lib()
    a, b = C.create_string_buffer(64), C.create_string_buffer(64)
    pa, pb = C.cast(a, C.c_void_p).value, C.cast(b, C.c_void_p).value
    hs = (C.c_void_p * 3)(pa, pb, pa)
    E = (C.c_double * 3)()
    rc = lib.dmrg_run_batch(hs, 3, 1, 2, C.cast(E, C.c_void_p), None)
    assert rc == dmrg_amd.ERR_BAD_ARG
    assert b"twice" in lib.dmrg_last_error()
