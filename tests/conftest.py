import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle check")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def J():
    import jayci_amd
    return jayci_amd


@pytest.fixture(scope="session")
def ctx(J):
    if J.device_count() < 1:
        pytest.fail("no GPU visible to a gpu-marked test")
    c = J.Context(0)
    yield c
    c.close()


def rand_vectors(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((k, n))


# small spaces covering FCI and truncated (docc/active/virtual) cases, both rules:
# (na, nb, norbs, ndocc, nactv, xlvl, rule)
SMALL_SPACES = [
    (2, 2, 4, 0, 4, 0, 1),      # FCI, 36 dets
    (3, 2, 6, 0, 6, 0, 1),      # FCI open shell, 300 dets
    (3, 3, 8, 0, 8, 0, 1),      # FCI 3a3b/8, 3136 dets
    (1, 0, 5, 0, 5, 0, 1),      # one electron: H = h
    (3, 3, 7, 1, 3, 2, 0),      # truncated, serial rule
    (3, 3, 8, 2, 2, 2, 1),      # truncated, MRCI-SD rule
    (4, 4, 10, 2, 4, 2, 1),     # MRCI-SD, virtual pairs
    (4, 3, 9, 1, 4, 1, 0),      # odd electrons, xlvl 1
]
