import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def qpch():
    """The C-ABI binding; GPU tests fail (not skip) when the library or device is missing."""
    from qoraphasecu_amd import _qpch

    assert _qpch.device_count() > 0, "no HIP device visible"
    return _qpch
