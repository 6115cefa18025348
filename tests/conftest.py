import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "cfd-art_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def have_gpu():
    try:
        import cfdart
        return cfdart.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    import cfdart
    if cfdart.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    return True
