"""Best-fit range bookkeeping of the device memory segments (host code),
compiled with g++ and checked against a per-unit occupancy model."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_range_set_best_fit_and_merge(tmp_path):
    exe = tmp_path / "ranges_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "digital_twin_brain-open_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ranges_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
