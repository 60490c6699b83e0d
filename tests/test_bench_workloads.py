"""CPU: the MPOs that bench.py builds for its workloads (its own copies, so
that the timed GPU legs import nothing from the oracle) are the oracle's
operators entry by entry, and so the pinned Kronecker Hamiltonians; the
workload table holds the configurations the results are quoted on."""
import numpy as np
import pytest

import bench
from oracle import models


@pytest.mark.parametrize("L", [2, 3, 6])
def test_heisenberg_mpo_is_the_oracle_mpo(L):
    Wb = bench.heisenberg_mpo(L)
    for a, b in zip(Wb, models.heisenberg_mpo(L), strict=True):
        assert np.array_equal(a, b)
    assert np.abs(models.mpo_to_dense(Wb) - models.dense_heisenberg(L).toarray()).max() == 0.0


def test_random_field_mpo_is_the_oracle_mpo():
    L = 6
    h = np.random.default_rng(3).uniform(-2, 2, L)
    h[2] = 0.0
    Wb = bench.heisenberg_mpo(L, h)
    for a, b in zip(Wb, models.random_field_heisenberg_mpo(L, h), strict=True):
        assert np.array_equal(a, b)
    Hk = models.dense_heisenberg(L, h=h).toarray()
    assert np.abs(models.mpo_to_dense(Wb) - Hk).max() < 1e-14


@pytest.mark.parametrize("n_max", [1, 2])
def test_josephson_mpo_is_the_oracle_mpo(n_max):
    kw = dict(n_max=n_max, E_J=1.3, E_C=0.7)
    Wb = bench.josephson_mpo(3, phi=0.4, **kw)
    for a, b in zip(Wb, models.josephson_mpo(3, phi_ext=0.4, **kw), strict=True):
        assert np.abs(a - b).max() == 0.0
    Hk = models.dense_josephson(3, phi_ext=0.4, **kw)
    Hk = Hk.toarray() if hasattr(Hk, "toarray") else Hk
    assert np.abs(models.mpo_to_dense(Wb) - Hk).max() < 1e-14


def test_workloads_are_the_stated_configurations():
    c4, c2, c3 = bench.WORKLOADS["c4"], bench.WORKLOADS["c2"], bench.WORKLOADS["c3"]
    assert (c4["model"], c4["L"], c4["chi"]) == ("heisenberg", 128, 512)
    assert (c2["model"], c2["L"], c2["chi"]) == ("heisenberg", 64, 256)
    assert (c3["model"], c3["L"], c3["chi"], c3["dtype"]) == ("josephson", 32, 256, "c128")
    Ws = bench.workload_mpo(c3)
    assert len(Ws) == 32 and Ws[0].shape == (c3["D"], c3["d"], c3["d"], c3["D"])
    assert np.iscomplexobj(Ws[0])


def test_batch_chains_are_seeded_and_uneven():
    wl = bench.WORKLOADS["c5"]
    chains = bench.batch_chains(wl)
    assert len(chains) == 64
    again = bench.batch_chains(wl)
    for (L, chi, W, h), (L2, chi2, W2, h2) in zip(chains, again):
        assert (L, chi, W) == (L2, chi2, W2) and np.array_equal(h, h2)
    for L, chi, W, h in chains:
        assert L in (16, 24, 32, 48, 64) and chi in (32, 64, 128, 256)
        assert h.shape == (L,) and np.all(np.abs(h) <= W)
    assert len({(L, chi) for L, chi, _, _ in chains}) > 8   # uneven sizes


def test_traffic_figure_comes_from_the_committed_profile():
    import json
    import os

    path = os.path.join(bench.ROOT, "profiles", "c4_heff_traffic.json")
    want = json.load(open(path))["bytes_per_matvec"]
    assert bench.heff_traffic("c4", "fp32") == want
    assert 1e8 < want < 1e9                       # the C4 centre-bond matvec: ~0.19 GB
    assert bench.heff_traffic("c4", "fp64") is None
    assert bench.heff_traffic("c2", "fp32") is None


def test_profile_summaries_name_their_commands():
    """Each rocprofv3 summary gives the command that made it, run on committed
    files (bench.py), never a scratch script."""
    import glob
    import os

    files = glob.glob(os.path.join(bench.ROOT, "profiles", "*.md"))
    assert files
    for f in files:
        text = open(f).read()
        assert "rocprofv3" in text and "bench.py" in text, f
        assert "_work" not in text, f


def test_cpu_baseline_bond_weighting():
    """The CPU baseline scales two timed centre bonds to a sweep by the
    bonds' H_eff flops: at C4 (L=128, chi=512) the 2*127 updates count as
    fewer centre-bond equivalents than 254 (the edge bonds are smaller) but
    more than the 2*110 full-size interior ones; a chain of equal bonds
    counts exactly 2(L-1)."""
    wl = bench.WORKLOADS["c4"]
    eq = bench.sweep_bond_equivalents(wl, wl["L"] // 2 - 1, 2)
    assert 2 * 110 < eq < 2 * 127
    flat = dict(wl, chi=1)
    assert abs(bench.sweep_bond_equivalents(flat, 60, 2) - 2 * (wl["L"] - 1)) < 1e-9
