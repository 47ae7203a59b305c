import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs through libmbx.so on cuda:0)")


@pytest.fixture(scope="session")
def lib():
    import merle_amd
    return merle_amd.load()


@pytest.fixture(scope="session")
def dev():
    from merle_amd.host import Device
    d = Device(0)
    yield d
    d.close()
