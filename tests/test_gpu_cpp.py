"""Runs the C++ drop-in suite (tests/cpp/test_hip_backend.cpp): HipBackend through the reference's
Backend interface, with the reference's exception types."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "test_hip_backend")


def test_cpp_backend_suite():
    assert os.path.exists(BIN), "build it first: make -C tests/cpp"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "0 failed, 0 skipped" in r.stderr
