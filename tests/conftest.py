import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqmgpu.so on the device)")


@pytest.fixture(scope="session")
def qm():
    """One device context for the GPU tests (tables are registered per test)."""
    from querymode_amd import QueryMode
    q = QueryMode.gpu(0)
    yield q
    q.close()
