import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "svo-tracer_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def svo():
    import _loader

    return _loader.load()


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(svo):
    if svo.device_count() < 1:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    cm = svo.ComputeManager(0)
    yield cm
    cm.Dispose()
