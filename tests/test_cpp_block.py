"""The C++ host class dtb::Block (include/dtb_block.hpp) over the C ABI:
compiles against the header and links libdtb.so; on a GPU it steps exactly
like the Python Block."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(dtb, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    lib = os.path.dirname(dtb.LIB_PATH)
    exe = tmp_path / "block_cpp"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "block_cpp.cpp"), "-L", lib, "-ldtb",
                    "-Wl,-rpath," + lib, "-o", str(exe)], check=True)
    return exe


def test_cpp_block_builds_and_reports_no_device(dtb, tmp_path):
    exe = _build(dtb, tmp_path)
    if dtb.device_count() > 0:
        pytest.skip("a GPU is visible; the GPU test covers the run")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("no device"), out.stdout + out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 3])   # DTB_LAYOUT_AUTO, DTB_LAYOUT_GATHER
def test_cpp_block_steps_like_python(dtb, gpu, tmp_path, layout):
    exe = _build(dtb, tmp_path)
    n, k, steps = 3000, 100, 200
    out = subprocess.run([str(exe), str(n), str(k), str(steps), str(layout)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    fields = out.stdout.split()
    assert fields[0] == "counts"
    step, events, spikes, total, vhash = (int(x) for x in fields[1:])

    net = dtb.gen_uniform(n, k, seed=1)
    blk = dtb.Block(net, delta_t=0.1, seed=11)
    try:
        blk.update_ou_background_stimuli(blk.subblk_id, 4.0, 1.5, 0.1)
        res = blk.run(steps, freqs=True)
        c = blk.counters()
        v = blk.state()["V"]
    finally:
        blk.engine.close()
    assert (step, events, spikes) == c and spikes > 0
    assert total == int(res["freq"].sum())
    assert vhash == int(np.bitwise_xor.reduce(v.view(np.uint32)))
