import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def hip():
    """The product module with a visible HIP device (gpu tests only)."""
    import ticit_amd
    props = ticit_amd.device_properties()
    if props["devices"] == 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return ticit_amd
