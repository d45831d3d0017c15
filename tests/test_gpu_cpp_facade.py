"""The reference device test-suite's cases through the C++ drop-in facade (tests/cpp/test_device_solver.cpp):
biofvm::solver -> solvers::device::solver -> C ABI -> HIP kernels."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parabiofvm_amd", "cpp", "build", "test_device_solver")


def test_cpp_device_suite():
    assert os.path.exists(EXE), "build the C++ facade first (__graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed cases" in r.stdout
