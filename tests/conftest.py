"""pytest set-up: the `gpu` marker, repository import path and shared helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmtmd.so")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.ORACLE_LIB):
        pyoracle.build(with_ref=False)
    return pyoracle
