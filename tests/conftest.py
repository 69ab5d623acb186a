import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def gpu_available():
    from icempm_3dm_amd import capi
    return capi.device_count() > 0


@pytest.fixture
def require_gpu(gpu_available):
    if not gpu_available:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
