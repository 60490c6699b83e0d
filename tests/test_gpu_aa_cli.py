"""The reference-compatible CLI (gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:474-536):
positional L chi_max n_sweeps, the model flags, the stdout contract
(dmrg2_gpu_impl.h:1451-1453: 'Final energy: %.12f', 'Total wall time: %.3f s',
'env_build_sec: ...') and the known-answer PASS lines of the reference's test
drivers (:208-211, :292-297, :344-346, :390-394, :442-447).

The binary runs as a child process.  This module is collected before every
other GPU test, so the pytest process itself has not touched the GPU yet."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dmrg-implementations_amd", "bin", "dmrg2_gpu_amd")


def run_cli(args):
    if not os.path.exists(BIN):
        raise FileNotFoundError("CLI not built: %s (run __graft_entry__.build())" % BIN)
    p = subprocess.run([BIN] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=110)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("args,E_ref", [
    (["8", "32", "10"], -3.374932598688),
    (["4", "32", "10", "--josephson"], -1.053346829927396),
    (["8", "32", "12", "--j1j2"], -3.0),
    (["8", "32", "12", "--j1j2j3", "--j2", "0.4", "--j3", "0.2"], -3.1205086380),
    (["8", "32", "12", "--j1j2j3", "--j2", "0.5", "--j3", "0.25"], -3.0690408638),
    (["4", "64", "12", "--ladder"], -4.2930664566945656),
])
def test_cli_known_answers(args, E_ref):
    rc, out, err = run_cli(args)
    assert rc == 0, out + err
    m = re.search(r"^Final energy: (-?\d+\.\d{12})$", out, re.M)
    assert m, out
    assert abs(float(m.group(1)) - E_ref) < 1e-8
    assert re.search(r"^Total wall time: \d+\.\d{3} s$", out, re.M), out
    assert re.search(r"^  env_build_sec: \d+\.\d{3}  timer_scope: sweep_only$", out, re.M), out
    assert re.search(r"^PASS$", out, re.M), out


def test_cli_quiet_keeps_result_lines():
    """The reference campaign harness always passes --quiet and parses the
    energy (benchmarks/run_campaign.py:262-278): the result lines stay."""
    rc, out, err = run_cli(["8", "32", "10", "--nmax", "1", "--quiet"])
    assert rc == 0, out + err
    m = re.search(r"^Final energy: (-?\d+\.\d{12})$", out, re.M)
    assert m and abs(float(m.group(1)) + 3.374932598688) < 1e-8, out
    assert "Total wall time:" in out and "env_build_sec:" in out
    assert "PASS" not in out
