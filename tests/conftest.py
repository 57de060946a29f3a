import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def device():
    """One HIP context for the whole GPU session (bound to torch's stream)."""
    from cyclix_msparc_amd.api import Device
    d = Device(0)
    yield d
    d.close()
