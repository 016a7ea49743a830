import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_vectors.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="session")
def queue():
    import compute_amd as ca
    q = ca.Queue(0)
    yield q
    q.close()


@pytest.fixture(autouse=True)
def _check_device_faults(request):
    """After every GPU test, finish the queue: this also reports a kernel that
    gave up a bounded look-back wait (the library raises instead of hanging)."""
    yield
    if request.node.get_closest_marker("gpu") is not None and "queue" in request.fixturenames:
        request.getfixturevalue("queue").finish()
