import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the engine through the C ABI")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o

    o.load()
    return o
