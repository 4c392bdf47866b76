import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def ctx():
    import torch

    import damascene_amd as dm

    assert torch.cuda.is_available(), "GPU test run without a GPU"
    c = dm.Context(0)
    yield c
    c.synchronize()
    c.close()
