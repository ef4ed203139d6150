import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def hg_ctx():
    """One GPU context for the session.  A missing libhg.so is an error, not a
    skip; only the absence of a visible GPU skips."""
    from libhybridgrains_amd import _lib

    _lib.lib()  # raises ImportError when the engine is not built
    try:
        ctx = _lib.Context(0)
    except _lib.HgError as e:
        pytest.skip(f"no GPU visible: {e}")
    yield ctx


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU work, so a crash is attributable
    if request.node.get_closest_marker("gpu"):
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
