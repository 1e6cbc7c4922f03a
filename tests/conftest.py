import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(autouse=True)
def fresh_omega_state():
    """Every test starts with the reference generator's cached variate as in
    a new process (the library keeps it process-wide, as the reference's
    statics are)."""
    try:
        from hqrrp_amd import _lib
        if os.path.exists(_lib.LIB_PATH):
            _lib.lib().hqrrp_omega_carry_set(0, 0.0)
    except (ImportError, OSError, AttributeError):
        pass
    yield


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
