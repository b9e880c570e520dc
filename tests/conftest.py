import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda:0")


def rel_l2(ax, ay, bx, by):
    import numpy as np
    num = np.sum((np.asarray(ax) - bx) ** 2 + (np.asarray(ay) - by) ** 2)
    den = np.sum(np.asarray(bx) ** 2 + np.asarray(by) ** 2)
    return float(np.sqrt(num / den))
