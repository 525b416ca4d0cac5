import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def ctx():
    """Device context for GPU tests.  Fails (does not skip) without a device:
    a GPU test must never pass on a host fallback."""
    import mossquantum_amd as mq
    n = mq.device_count()
    assert n >= 1, "no HIP device visible: GPU tests need an MI355X"
    c = mq.default_context()
    yield c


def pytest_collection_modifyitems(config, items):
    # print the case name before each GPU test so a device fault names its test
    for item in items:
        if item.get_closest_marker("gpu"):
            item.user_properties.append(("gpu", True))


@pytest.fixture(autouse=True)
def _announce(request):
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield
