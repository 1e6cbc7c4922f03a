import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def handle():
    import torch

    import hqrrp_amd

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = hqrrp_amd.Handle(0)
    h.use_stream(torch.cuda.current_stream())
    yield h
    h.close()
