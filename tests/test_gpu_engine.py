"""End-to-end engine runs on the GPU against exact diagonalisation and the
reference's known answers (gpu-rocm/dmrg2-gpu/src/test_dmrg2_gpu.cpp:208-297)."""
import numpy as np
import pytest

from oracle import dmrg as od
from oracle import models

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", [2, 1])
@pytest.mark.parametrize("L,E_ref", [(4, -1.616025403784), (8, -3.374932598688)])
def test_heisenberg_known_answers(dmrg_amd, variant, L, E_ref):
    eng = dmrg_amd.DMRG(L, 2, 32, 5, variant=variant)
    eng.set_mpo(models.heisenberg_mpo(L))
    if variant == 2:
        eng.initialize_mps_neel()
    else:
        eng.initialize_mps_random(seed=4, scale=0.1)
    E = eng.run(30)
    assert abs(E - E_ref) < 1e-10
    eng.close()


@pytest.mark.parametrize("L,E_ref", [(4, -1.053346829927396), (6, -1.748843818181493)])
def test_josephson_known_answers(dmrg_amd, L, E_ref):
    Ws = models.josephson_mpo(L, n_max=1, E_J=1.0, E_C=0.5, phi_ext=np.pi / 4)
    eng = dmrg_amd.DMRG(L, 3, 32, 4, dtype="c128")
    eng.set_mpo(Ws)
    eng.initialize_mps_random(seed=9, scale=0.1)
    E = eng.run(20)
    assert abs(E - E_ref) < 1e-8
    eng.close()


def test_two_site_matches_oracle_sweep_by_sweep(dmrg_amd):
    L, chi = 10, 16
    Ws = models.heisenberg_mpo(L)
    mps, b = od.random_mps(L, 2, chi, seed=21, scale=0.1)
    E_o, _, b_o, sw_o = od.dmrg2(Ws, mps, b, chi, 4, krylov_max=20)
    eng = dmrg_amd.DMRG(L, 2, chi, 5, krylov_max=20)
    eng.set_mpo(Ws)
    eng.initialize_mps_random(seed=21, scale=0.1)
    got0 = eng.get_mps()
    for A, B in zip(got0, mps):
        assert np.array_equal(A, B)  # bit-identical seeded start
    E = eng.run(4)
    sw = eng.sweep_energies()
    assert len(sw) == len(sw_o)
    assert np.max(np.abs(np.array(sw) - np.array(sw_o))) < 1e-10 * L
    E_ed = models.ground_energy(models.dense_heisenberg(L))
    assert E >= E_ed - 1e-10 and E - E_ed < 1e-7  # chi=16 < 32 truncates: variational
    e, n2 = eng.expectation()
    assert abs(n2 - 1) < 1e-10 and abs(e - E) < 1e-10 * L
    eng.close()


def test_fp32_matvec_mode_tracks_fp64(dmrg_amd):
    """FP32 H_eff matvecs with FP64 Lanczos/SVD/environments: the reported
    energy (FP64 Rayleigh quotient at the last bond) stays within 1e-9 per
    site of the FP64 run from the same seeded state."""
    L, chi = 16, 32
    Ws = models.heisenberg_mpo(L)
    out = {}
    for prec in (0, 1):
        eng = dmrg_amd.DMRG(L, 2, chi, 5, tol=0.0, krylov_max=20, precision=prec)
        eng.set_mpo(Ws)
        eng.initialize_mps_random(seed=7, scale=0.1)
        out[prec] = eng.run(6)
        e, n2 = eng.expectation()
        assert abs(n2 - 1) < 1e-9
        assert abs(e - out[prec]) < 1e-9 * L
        eng.close()
    assert abs(out[1] - out[0]) < 1e-9 * L, out
