"""CPU checks of the serving engine's host logic (no GPU): the arena
allocator (no overlap, exact accounting, full coalescing) and the request
loop (arrival-order admission, slot and budget limits, resident factors,
refusals), built from hpcc_fpga_amd/csrc/serve_core.hpp with g++ and driven by
a stand-in executor (tests/serve_core_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_serve_core_host_logic(tmp_path):
    exe = tmp_path / "serve_core_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-Wall", "-Werror",
                    "-fsanitize=address,undefined", os.path.join(HERE, "serve_core_check.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "OK"
