"""The k-space kernels' sin/cos (rbmd_amd/csrc/sincos.hpp), compiled for the
host from the same header, against long double sin/cos: FP64 within 2.5e-16,
FP32 within 1.2e-7 (absolute, on unit-amplitude values)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_sincos_turns_accuracy(tmp_path):
    exe = tmp_path / "sincos_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "rbmd_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "sincos_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    e64, e32 = float(out[0]), float(out[1])
    assert e64 < 2.5e-16, e64
    assert e32 < 1.2e-7, e32
