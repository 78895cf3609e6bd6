import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libjp2lam_amd.so")


@pytest.fixture(scope="session")
def ora():
    from oracle_ctypes import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def lib():
    from jp2lam_amd import _lib

    return _lib.load()


@pytest.fixture(scope="session")
def gpu(lib):
    if not lib.jp2g_device_available():
        pytest.fail("no HIP device: GPU tests need an MI355X")
    return lib
