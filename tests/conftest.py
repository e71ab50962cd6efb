import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def host_ctx():
    import vision_amd as va
    ctx = va.Context(-1)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def gpu_ctx():
    import vision_amd as va
    ctx = va.Context(0)
    yield ctx
    ctx.close()
