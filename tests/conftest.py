import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm device (MI355X) and the built libglide_hip.so")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O
