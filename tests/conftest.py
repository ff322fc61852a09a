import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    """The oracle (CPU checker), built on first use if needed."""
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ref_vectors():
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "tmv_ref_vectors.npz")
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    import tmv_amd
    tmv_amd.lib()          # fail loudly if the HIP library is missing
    return torch.device("cuda:0")
