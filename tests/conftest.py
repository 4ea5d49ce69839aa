import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built engine library")


@pytest.fixture(scope="session")
def engine_lib():
    from squishy_volumes_amd import abi

    lib = abi.load_library()  # raises when the HIP library is missing: no fallback
    if lib.mpm_device_count() < 1:
        pytest.fail("no HIP device visible to the engine")
    return lib
