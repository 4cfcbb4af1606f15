"""The stub build of the HIP factory (host/hip_stub.cpp, the counterpart of the reference's
src/cuda_stub.cpp): linked without liblettuce_hip.so, make_hip_backend returns nullptr plus why,
hip_compiled() is false and hip_device_count() is 0. Runs on the CPU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "bin", "test_hip_stub")


def test_stub_build_links_without_the_hip_library_and_reports_why():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "bin/test_hip_stub"], check=True)
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "lettuce_hip" not in out and "amdhip" not in out
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "no HIP support" in r.stderr and "0 failed" in r.stderr
