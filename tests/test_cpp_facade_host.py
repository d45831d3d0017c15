"""Host-only cases of the C++ drop-in facade (tests/cpp/test_device_solver.cpp, names starting host_): the
builder's bookkeeping and the agent container's swap-remove slots, with no device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "parabiofvm_amd", "cpp", "build", "test_device_solver")


def test_cpp_facade_host_cases():
    if not os.path.exists(EXE):
        pytest.skip("C++ facade not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, "host_"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "2 cases" in r.stdout and "0 failed cases" in r.stdout, r.stdout
