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
