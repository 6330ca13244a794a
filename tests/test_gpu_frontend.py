"""Runs the C++ front-end test program (tests/cpp/test_frontend.cpp, built by build())
on the GPU: heffte::fft3d / fft3d_r2c / rtransform with backend::mi355 tags, vector and
array APIs, workspace, batch, in place, exceptions."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_frontend.bin")


@pytest.mark.gpu
def test_cpp_frontend():
    assert os.path.exists(BIN), "build() compiles tests/cpp/test_frontend.bin"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
