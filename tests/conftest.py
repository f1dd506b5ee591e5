import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "hifiles-solver_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available():
    """True when a HIP device is visible (without initialising it through torch)."""
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "":
        return False
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test run without a visible GPU (/dev/kfd missing)")
