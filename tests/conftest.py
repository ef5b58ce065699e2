import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def dtb():
    return importlib.import_module("digital_twin_brain-open_amd")


@pytest.fixture(scope="session")
def orc():
    return importlib.import_module("oracle")


@pytest.fixture(scope="session")
def gpu(dtb):
    if dtb.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return 0
