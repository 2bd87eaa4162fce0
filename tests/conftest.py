import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO / "mose-gpu_amd", REPO / "tests", REPO):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def _ensure_built():
    from mosegpu.core import LIB_PATH
    if not LIB_PATH.exists():
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def gpu_ready():
    _ensure_built()
    import mosegpu
    n = mosegpu.device_count()
    if n < 1:
        pytest.fail("no GPU visible: the gpu-marked tests run on an MI355X")
    return n
