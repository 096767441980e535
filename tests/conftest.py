import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_report_header(config):
    return f"rta_amd tests, root {ROOT}"


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a HIP device is visible; initialises device 0."""
    import rta_amd
    if rta_amd.device_count() < 1:
        pytest.skip("no HIP device")
    rta_amd.call("rta_hip_set_device", 0)
    return rta_amd.device_name()
