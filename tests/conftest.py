"""Test configuration.

Markers: `gpu` tests need a HIP device and call the C ABI of librgs_hip.so;
everything else runs on the CPU (the oracle, the host logic, and a load check
of the library).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built librgs_hip.so")
    config.addinivalue_line("markers", "slow: long-running CPU check, excluded from the default CPU suite")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip_ctx():
    """A library context on device 0 (GPU tests only)."""
    from rustscan_amd import build
    from rustscan_amd.rasterizer import Context

    if not os.path.exists(build.LIB_PATH):
        build.build_hip()
    ctx = Context(0)
    yield ctx
    ctx.close()
