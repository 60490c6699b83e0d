import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dmrg-implementations_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


@pytest.fixture(scope="session")
def dmrg_amd():
    import dmrg_amd as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def ops(dmrg_amd):
    o = dmrg_amd.Ops(0)
    yield o
    o.close()


def rand(rng, shape, cplx=False):
    a = rng.uniform(-1, 1, size=shape)
    if cplx:
        a = a + 1j * rng.uniform(-1, 1, size=shape)
    return a


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
