import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def gpu_available():
    try:
        import tinypt_amd

        return tinypt_amd.hip_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    if not gpu_available():
        pytest.skip("no HIP device visible")
    import tinypt_amd

    return tinypt_amd
