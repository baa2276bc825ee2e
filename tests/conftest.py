import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def _gpu_present() -> bool:
    return os.path.exists("/dev/kfd") and os.path.exists("/dev/dri")


@pytest.fixture(scope="session")
def ctx():
    """One libfsx context for the whole GPU session (one process, one GPU)."""
    if not _gpu_present():
        pytest.fail("GPU test selected but no GPU is visible")
    from flecsolve_amd import fsx
    c = fsx.Context(0)
    yield c
    c.close()
