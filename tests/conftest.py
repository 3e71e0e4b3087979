import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def gpu_available() -> bool:
    try:
        import simplehydrology_amd as shy
        return shy.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def shy():
    import simplehydrology_amd as m
    if m.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle
