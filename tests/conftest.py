import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the reference's sources compiled in place)")


@pytest.fixture(scope="session")
def gm():
    """One initialised GPU module per test session (fails loudly without a device)."""
    from simstack_amd import gpu
    return gpu.GpuModule(arena_bytes=4 << 30)
