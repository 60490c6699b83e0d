"""Batched independent chains (dmrg_run_batch): random-field Heisenberg
chains of uneven L and chi, FP32-matvec mode, run concurrently; each chain's
energy is checked against its own sequential FP64 solve from the same seeded
state and against exact diagonalisation."""
import numpy as np
import pytest

from oracle import models

pytestmark = pytest.mark.gpu


def _chains():
    rng = np.random.default_rng(5)
    out = []
    for L, chi, W in [(8, 16, 0.5), (12, 32, 2.0), (10, 16, 1.0), (12, 16, 5.0), (8, 32, 2.0),
                      (10, 32, 0.5)]:
        h = rng.uniform(-W, W, L)
        out.append((L, chi, h))
    return out


def test_batch_matches_sequential_fp64(dmrg_amd):
    chains = _chains()
    sweeps = 8
    ref = []
    for i, (L, chi, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=30)
        e.set_mpo(models.random_field_heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=100 + i, scale=0.1)
        ref.append(e.run(sweeps))
        e.close()
    engines = []
    for i, (L, chi, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=30, precision=dmrg_amd.PREC_FP32)
        e.set_mpo(models.random_field_heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=100 + i, scale=0.1)
        engines.append(e)
    E, T = dmrg_amd.run_batch(engines, sweeps, threads=3)
    for e in engines:
        e.close()
    for i, (L, chi, h) in enumerate(chains):
        assert abs(E[i] - ref[i]) < 1e-9 * L, (i, E[i], ref[i])
        E_ed = models.ground_energy(models.dense_heisenberg(L, h=h))
        assert E[i] >= E_ed - 1e-9 and E[i] - E_ed < 1e-5, (i, E[i], E_ed)
        assert T[i] > 0


def test_batch_reports_chain_error(dmrg_amd):
    # a chain without an MPO fails with the STATE code; the batch returns it
    ok = dmrg_amd.DMRG(6, 2, 8, 5)
    ok.set_mpo(models.heisenberg_mpo(6))
    ok.initialize_mps_random(seed=1)
    bad = dmrg_amd.DMRG(6, 2, 8, 5)
    with pytest.raises(dmrg_amd.DMRGError) as ei:
        dmrg_amd.run_batch([ok, bad], 2, threads=2)
    assert ei.value.code == 5
    ok.close()
    bad.close()


def test_batch_streams_returned_to_engines(dmrg_amd):
    """The batch lends each worker's stream to the chains it runs and hands
    the engines their own streams back: continuing every chain on its own
    after a batch of 3 sweeps, then a second batch of 2, gives each chain
    exactly the energy of 5 + 2 sweeps run one chain at a time (the kernels
    and their order are the same; only the streams differ)."""
    chains = _chains()
    ref = []
    for i, (L, chi, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=30)
        e.set_mpo(models.random_field_heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=200 + i, scale=0.1)
        e.run(3)
        e.run(2)
        ref.append(e.run(2))
        e.close()
    engines = []
    for i, (L, chi, h) in enumerate(chains):
        e = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=30)
        e.set_mpo(models.random_field_heisenberg_mpo(L, h))
        e.initialize_mps_random(seed=200 + i, scale=0.1)
        engines.append(e)
    dmrg_amd.run_batch(engines, 3, threads=2)
    mid = [e.run(2) for e in engines]  # on the engines' own streams again
    E, _ = dmrg_amd.run_batch(engines, 2, threads=4)
    for e in engines:
        e.close()
    assert all(np.isfinite(mid))
    assert list(E) == ref, (list(E), ref)
