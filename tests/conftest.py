import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, initialised on a real device; fails loudly otherwise."""
    import dust2_amd
    from dust2_amd._lib import call
    dust2_amd.lib()
    call("dust2_amd_init")
    return dust2_amd
