import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fairsim-vigor_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfsim.so on cuda:0)")


def _step(msg):
    print(f"[fsim-test] {msg}", flush=True)


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu"):
        _step(f"start {request.node.nodeid}")
    yield


@pytest.fixture(scope="session")
def fsim_lib():
    from fsim import _lib
    _lib.init(0)
    return _lib
