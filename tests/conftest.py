"""Test configuration: registers the gpu marker and puts the package mirror
(ufo-ir_amd/uir.py) and the oracle binding (oracle/oracle.py) on sys.path."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ufo-ir_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libuir.so")


@pytest.fixture(scope="session")
def res():
    import uir
    if uir.device_count() < 1:
        pytest.skip("no HIP device visible")
    r = uir.Resources(0)
    yield r
    r.close()
