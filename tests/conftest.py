import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (test infrastructure)."""
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def nl():
    """The GPU library's Python host layer; fails loudly without a GPU build."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import nalgebra_amd
    return nalgebra_amd
