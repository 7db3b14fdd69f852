import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available() -> bool:
    try:
        import craystal_amd as cr
        return cr.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a gpu-marked test on a box without a GPU must fail
    loudly, not fall back to anything."""
    import craystal_amd as cr
    cr.load()
    assert cr.device_count() > 0, "no GPU visible to libcrayhip.so"
    return cr.get_context(0)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
