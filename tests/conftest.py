import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def native():
    """The product library; built on demand when absent (CPU-only build is fine)."""
    from openems_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        from openems_amd import build
        build.build()
    return _native.lib()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    return oracle.lib()


@pytest.fixture(scope="session")
def gpu(native):
    if native.ems_device_count() <= 0:
        pytest.fail("no HIP device visible: GPU tests must run on the GPU host")
    return native
