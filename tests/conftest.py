import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP kernels through the C ABI")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def hip():
    """The device library, once a HIP device is present."""
    import simcraft_amd

    if simcraft_amd.device_count() == 0:
        pytest.skip("no HIP device")
    return simcraft_amd


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding

    binding.load()
    return binding


def announce(msg: str) -> None:
    """Print the case name before a GPU step so a failing run names it."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
