import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "clay-engine_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def root():
    """One HIP context for the whole GPU session."""
    import clayhip

    r = clayhip.Root(0)
    yield r
    r.destroy()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    O.set_threads(min(16, os.cpu_count() or 1))
    return O
