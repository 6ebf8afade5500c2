import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pytest_collection_modifyitems(config, items):
    # GPU tests are collected everywhere; without a device they fail loudly
    # rather than silently passing on a fallback.
    pass


@pytest.fixture(scope="session")
def cpt():
    import cudapt_amd
    cudapt_amd.lib()
    if cudapt_amd.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return cudapt_amd
