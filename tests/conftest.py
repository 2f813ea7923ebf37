import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libefl.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One efl_ctx for the whole GPU session (one process on the card)."""
    from bevy_eulerian_fluid_amd import native
    ctx = native.Context(0)
    yield ctx
    ctx.close()
