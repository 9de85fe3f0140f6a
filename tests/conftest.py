import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def lib():
    """The ampcor_hip module on a machine with a GPU; fails loudly when missing."""
    from ampcor_amd import ext

    mod = ext.require()
    assert mod.device_count() > 0, "no HIP device visible to the HIP library"
    return mod
