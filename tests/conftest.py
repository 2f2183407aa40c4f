import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP library")


@pytest.fixture(scope="session")
def pa():
    """The GPU library, initialised on device 0 (GPU tests only)."""
    import polymer_amd
    polymer_amd.init(0)
    print("device:", polymer_amd.device_name(), flush=True)
    return polymer_amd


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    return pyoracle
