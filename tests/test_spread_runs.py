"""The spread kernels find a tile's candidate points as runs of the
bin-sorted points (spread_tiles.h: tile_bins, tile_run) and walk them with a
per-lane cursor.  This compiles the real header for the host with hipcc and
checks, for every tile of several geometries, that the cursor stays inside
the run table, every run position inside the sorted points, and that the
runs visit exactly the points whose kernel stencil reaches the tile (each
once)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tile_runs_cover_exactly(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not installed")
    exe = str(tmp_path / "runs_check")
    subprocess.check_call([hipcc, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "mp_tools_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "runs_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
