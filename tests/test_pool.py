"""The runtime's device-memory arena (sister_amd/csrc/pool.h), driven with host
memory by tests/native/pool_check.cpp: alignment, disjoint blocks, the budget,
coalescing of returned blocks and trim()."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pool_arena(tmp_path):
    exe = str(tmp_path / "pool_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", os.path.join(ROOT, "tests", "native", "pool_check.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
