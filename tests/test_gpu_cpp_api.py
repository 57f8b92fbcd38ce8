"""Backend::hip through the public C++ headers: runs tests/cpp/test_hip_backend (built by
qoraphasecu_amd/build.py), the reference's C++-only solver checks (progress schedule and
values, a run continuing from the current state bit for bit, ER monotone, argument
errors) for the HIP backend."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_hip_backend")


def test_cpp_api_cases():
    assert os.path.exists(BIN), "tests/cpp/test_hip_backend is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 case(s) failed" in r.stdout
