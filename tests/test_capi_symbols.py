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


def _header_prototypes():
    """name -> (return type, [parameter types]) from include/dmrg_amd.h."""
    hdr = open(os.path.join(ROOT, "include", "dmrg_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", " ", hdr, flags=re.S)
    protos = {}
    for m in re.finditer(r"(const char\s*\*|int|void)\s+(dmrg_\w+)\s*\(([^)]*)\)\s*;", hdr):
        params = [p.strip() for p in m.group(3).split(",") if p.strip() and p.strip() != "void"]
        types = [re.sub(r"\s*\b\w+$", "", p) if not p.endswith("*") else p for p in params]
        protos[m.group(2)] = (m.group(1).replace(" ", ""), types)
    return protos


def _kind_c(t):
    t = t.replace("const ", "").strip()
    if "*" in t or t in ("dmrg_handle", "dmrg_ctx_handle"):
        return "ptr"
    return {"int": "int", "double": "double", "uint64_t": "u64", "size_t": "u64"}[t]


def _kind_ctypes(t):
    import ctypes as C
    if t is None:
        return "void"
    if t in (C.c_void_p, C.c_char_p) or (isinstance(t, type) and issubclass(t, C._Pointer)):
        return "ptr"
    if t is C.c_int:
        return "int"
    if t is C.c_double:
        return "double"
    if t in (C.c_uint64, C.c_size_t):
        return "u64"
    raise AssertionError("unexpected ctypes type %r" % (t,))


def test_ctypes_signatures_match_the_header(dmrg_amd):
    """Every prototype of the header against the binding's ctypes table:
    argument count and kind (pointer, int, double, 64-bit unsigned), and the
    return kind, so a drifted signature cannot pass arguments silently wrong."""
    lib = dmrg_amd.lib()
    protos = _header_prototypes()
    assert set(dmrg_amd.EXPORTED) <= set(protos)
    for name, (ret, params) in protos.items():
        f = getattr(lib, name)
        assert f.argtypes is not None, name
        got = [_kind_ctypes(t) for t in f.argtypes]
        assert got == [_kind_c(p) for p in params], (name, got, params)
        want_ret = {"int": "int", "void": "void", "constchar*": "ptr"}[ret]
        assert _kind_ctypes(f.restype) == want_ret, name


def test_missing_library_raises(dmrg_amd, monkeypatch, tmp_path):
    """No fallback: without the built library the binding raises."""
    import pytest

    monkeypatch.setattr(dmrg_amd, "_lib", None)
    monkeypatch.setattr(dmrg_amd, "LIB_PATH", str(tmp_path / "libdmrg_amd.so"))
    with pytest.raises(ImportError):
        dmrg_amd.lib()


def test_every_entry_point_cites_the_reference_or_says_it_has_none():
    """Each prototype's comment cites the reference interface it replaces
    (file:line) or states that it has no counterpart.  A prototype that
    directly follows another, with no comment of its own, shares that one's."""
    hdr = open(os.path.join(ROOT, "include", "dmrg_amd.h")).read()
    prev, last, missing, n = 0, "", [], 0
    for m in re.finditer(r"(const char\s*\*|int|void)\s+(dmrg_\w+)\s*\(([^)]*)\)\s*;", hdr):
        comments = re.findall(r"/\*.*?\*/", hdr[prev:m.start()], flags=re.S)
        c = comments[-1] if comments else last
        if not (re.search(r"\.(h|hpp|cpp|py):\d+", c) or "counterpart" in c.lower()):
            missing.append(m.group(2))
        last, prev, n = c, m.end(), n + 1
    assert n == len(_header_prototypes())
    assert not missing, missing


def test_integration_guide_names_every_entry_point():
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    missing = [f for f in sorted(_header_prototypes()) if f not in guide]
    assert not missing, missing
