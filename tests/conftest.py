import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    binding.lib()
    return binding


@pytest.fixture(scope="session")
def device():
    from malevich_amd import pipeline

    if pipeline.device_count() < 1:
        pytest.fail("GPU test selected but no HIP device is visible")
    dev = pipeline.Device(0)
    yield dev
    dev.close()
