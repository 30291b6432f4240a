import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running")


def _ensure_lib():
    from inching_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from inching_amd.build import build
        build(verbose=False)
    return _lib.lib()


@pytest.fixture(scope="session")
def lib():
    return _ensure_lib()


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a GPU (run with -m 'not gpu')")
    _ensure_lib()
    return torch.device("cuda", 0)
