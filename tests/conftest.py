"""Test configuration: the `gpu` marker, the repository on sys.path, and shared fixtures."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


@pytest.fixture(scope="session")
def device():
    """The GPU, opened once per session. There is no CPU fallback: without a GPU this raises."""
    import libsci_amd

    print(f"[tests] opening GPU 0", flush=True)
    d = libsci_amd.Device(0)
    print(f"[tests] device: {d.name}", flush=True)
    yield d
    d.close()
