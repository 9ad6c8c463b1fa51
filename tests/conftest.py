import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def si8():
    from rsdft_amd.system import make_config
    return make_config("si8")


@pytest.fixture(scope="session")
def oracle_si8(si8):
    from oracle.oracle import Oracle
    return Oracle(si8)


@pytest.fixture(scope="session")
def gpu_ctx():
    """Loads the HIP library; fails loudly when it is missing."""
    from rsdft_amd import gpu
    return gpu.load_library()
