"""The kernels replace integer division by a runtime divisor with a
multiply-shift (mpsq_device.h: make_fdiv / fdiv).  This compiles the real
header for the host with hipcc and checks it against integer division
exhaustively for small divisors and at the edges for large ones."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fdiv_exact(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not installed")
    exe = str(tmp_path / "fdiv_check")
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mp_tools_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "fdiv_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
