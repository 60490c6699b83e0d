"""Argument and input errors of the reference-compatible CLI, which are
reported before the engine touches a GPU (the reference driver exits with 1
on an unknown flag: gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:510)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dmrg-implementations_amd", "bin", "dmrg2_gpu_amd")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(args):
    if not os.path.exists(BIN):
        pytest.fail("CLI not built: %s (run __graft_entry__.build())" % BIN)
    p = subprocess.run([BIN] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=60)
    return p.returncode, p.stdout, p.stderr


def test_unknown_flag():
    rc, out, err = run(["8", "32", "2", "--no-such-flag"])
    assert rc == 1 and "Unknown flag: --no-such-flag" in err


def test_missing_flag_value():
    rc, out, err = run(["8", "32", "2", "--ej"])
    assert rc == 1 and "missing value for --ej" in err


def test_mps_needs_mpo():
    rc, out, err = run(["12", "10", "2", "--mps", os.path.join(GOLD, "heisenberg_L12_chi10_mps.bin")])
    assert rc == 1 and "--mps needs --mpo" in err


def test_truncated_bin_rejected(tmp_path):
    raw = open(os.path.join(GOLD, "heisenberg_L12_mpo.bin"), "rb").read()
    bad = tmp_path / "trunc.bin"
    bad.write_bytes(raw[: len(raw) - 16])
    rc, out, err = run(["12", "10", "2", "--mpo", str(bad)])
    assert rc == 1 and err.startswith("ERROR:")
    assert "Final energy" not in out
