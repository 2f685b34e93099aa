import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import globe_amd

    if globe_amd.device_count() < 1:
        pytest.fail("gpu test collected but no GPU is visible")

    def make():
        return globe_amd.Globe(0)

    return make
