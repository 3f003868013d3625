"""Run the C++ drop-in test program (tests/cpp/test_ffts_cpp.cpp) on the GPU."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_ffts_cpp.bin")


def test_cpp_dropin_program(gpu):
    assert os.path.exists(BIN), "build the C++ tests first (__graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "ALL PASSED" in r.stdout
