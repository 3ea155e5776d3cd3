import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "mlx-learn_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture
def rng():
    # seed of the reference suite (tests/conftest.py:36-38)
    return np.random.default_rng(20260726)


@pytest.fixture(autouse=True)
def _reset_hooks():
    yield
    try:
        import torch
        if torch.cuda.is_available():
            from mlxlearn_amd import _lib
            _lib.set_test_hooks(0, 0)
            _lib.set_tuning()
    except Exception:
        pass
