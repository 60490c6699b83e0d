"""CPU: the MPOs that the CLI hands to dmrg_set_mpo (its own C++ builders in
csrc/cli.cpp, and its reading of the reference's .bin files) are the oracle's
pinned operators.  `--dump-mpo FILE` writes them before any device call.

Energies alone cannot pin a builder: a hopping sign flipped once per term, or
a conjugated flux phase, leaves the spectrum unchanged.  So each CLI model is
compared with the oracle's Kronecker Hamiltonian entry by entry, and with the
oracle's MPO where the two builders share a gauge."""
import os
import subprocess

import numpy as np
import pytest

from oracle import models

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dmrg-implementations_amd", "bin", "dmrg2_gpu_amd")
GOLD = os.path.join(ROOT, "tests", "golden")


def dump(tmp_path, args):
    """Run the CLI with --dump-mpo and return the site tensors W[w, s, sp, wp]."""
    if not os.path.exists(BIN):
        pytest.fail("CLI not built: %s (run __graft_entry__.build())" % BIN)
    out = str(tmp_path / "mpo.dump")
    p = subprocess.run([BIN] + args + ["--dump-mpo", out], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0, p.stderr
    raw = open(out, "rb").read()
    L, D, d, cx = np.frombuffer(raw[:32], dtype=np.int64)
    data = np.frombuffer(raw[32:], dtype=np.complex128 if cx else np.float64)
    n = D * d * d * D
    assert data.size == L * n
    return [data[i * n:(i + 1) * n].reshape((D, d, d, D), order="F") for i in range(L)]


def dense(H):
    return H.toarray() if hasattr(H, "toarray") else np.asarray(H)


def test_heisenberg(tmp_path):
    Ws = dump(tmp_path, ["5", "8", "1"])
    assert np.abs(models.mpo_to_dense(Ws) - dense(models.dense_heisenberg(5))).max() == 0.0
    for a, b in zip(Ws, models.heisenberg_mpo(5), strict=True):
        assert np.array_equal(a, b)


def test_tfim(tmp_path):
    Ws = dump(tmp_path, ["5", "8", "1", "--tfim", "--hfield", "0.7"])
    assert np.abs(models.mpo_to_dense(Ws) - dense(models.dense_tfim(5, J=1.0, hx=0.7))).max() < 1e-15


@pytest.mark.parametrize("flag,Js", [("--j1j2", (1.0, 0.45)), ("--j1j2j3", (1.0, 0.4, 0.2))])
def test_frustrated_chains(tmp_path, flag, Js):
    args = ["6", "8", "1", flag, "--j1", str(Js[0]), "--j2", str(Js[1])]
    if len(Js) == 3:
        args += ["--j3", str(Js[2])]
    Ws = dump(tmp_path, args)
    H = dense(models.dense_spin_chain(6, list(Js)))
    assert np.abs(models.mpo_to_dense(Ws) - H).max() < 1e-14


def test_ladder(tmp_path):
    Ws = dump(tmp_path, ["3", "8", "1", "--ladder", "--jleg", "1.0", "--jrung", "0.7"])
    H = dense(models.dense_ladder(3, J_leg=1.0, J_rung=0.7))
    assert np.abs(models.mpo_to_dense(Ws) - H).max() < 1e-14


@pytest.mark.parametrize("n_max", [1, 2])
def test_josephson(tmp_path, n_max):
    Ws = dump(tmp_path, ["3", "8", "1", "--josephson", "--nmax", str(n_max), "--ej", "1.3",
                         "--ec", "0.7", "--phi", "0.4"])
    kw = dict(n_max=n_max, E_J=1.3, E_C=0.7, phi_ext=0.4)
    assert np.abs(models.mpo_to_dense(Ws) - dense(models.dense_josephson(3, **kw))).max() < 1e-14
    for a, b in zip(Ws, models.josephson_mpo(3, **kw), strict=True):
        assert np.abs(a - b).max() < 1e-15


@pytest.mark.parametrize("stem,mps", [("heisenberg_L12", "heisenberg_L12_chi10_mps.bin"),
                                      ("josephson_L8_n2", "josephson_L8_n2_chi10_mps.bin")])
def test_reference_bin_files(tmp_path, io, stem, mps):
    """The CLI's reading of the reference's .bin MPO equals the Python importer's."""
    mpo = os.path.join(GOLD, stem + "_mpo.bin")
    Ws = dump(tmp_path, ["--mpo", mpo])
    Wp, _, _ = io.load_bin_case(os.path.join(GOLD, mps), mpo)
    for a, b in zip(Ws, Wp, strict=True):
        assert a.shape == np.shape(b)
        assert np.abs(a - b).max() == 0.0


@pytest.fixture
def io():
    from dmrg_amd import io as m
    return m
