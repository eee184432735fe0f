import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SEC_PER_YEAR = 31557600.0


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs librsfg kernels)")


def _gpu_count():
    from interact_amd import lib
    return lib.device_count()


def pytest_collection_modifyitems(config, items):
    # GPU tests load the library (fail loudly if it is missing) and only skip
    # when no device is visible (CPU-only containers)
    if any("gpu" in it.keywords for it in items):
        try:
            ngpu = _gpu_count()
        except Exception:
            ngpu = -1
        if ngpu == 0:
            skip = pytest.mark.skip(reason="no GPU visible")
            for it in items:
                if "gpu" in it.keywords:
                    it.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
