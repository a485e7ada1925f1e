"""Shared test setup: the `gpu` marker and a session-wide HIP device."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: takes more than ~20 s on the CPU")


@pytest.fixture(scope="session")
def device():
    """The HIP device. Fails loudly (no fallback) if the library or GPU is absent."""
    import kutta_amd
    dev = kutta_amd.Device(0)
    print(f"[device] {dev.name}", flush=True)
    yield dev
    dev.close()
