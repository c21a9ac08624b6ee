"""Runs the C++ restatement of the reference's ivp tests (tests/cpp/
test_reference_ivp.cpp) against the C++ builder mirror include/bacon_ivp.hpp
and libbacon_gpu.so on the GPU.  The binary is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_reference_ivp.bin")


def test_cpp_reference_ivp_tests(gpu):
    assert os.path.exists(BIN), "build the C++ test with `python -m bacon_amd.build`"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK (0 failures)")
