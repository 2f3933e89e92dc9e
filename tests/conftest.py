import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and calls the C ABI")


@pytest.fixture(scope="session")
def orc():
    from tests import oracle_lib

    return oracle_lib.Oracle()


@pytest.fixture(scope="session")
def rp():
    import rypath_amd

    return rypath_amd


@pytest.fixture(scope="session")
def gpu_available(rp):
    if rp.device_count() <= 0:
        pytest.skip("no HIP device")
    return True
