"""CPU test of the single-block entry points' per-thread context setup: built
with g++ from icbc_amd/csrc/block_context.h and a fake runtime that fails at
each step in turn (a failed setup must leave no half-made context behind)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_block_context_setup_is_all_or_nothing(tmp_path):
    exe = tmp_path / "block_context_test"
    subprocess.check_call(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "icbc_amd", "csrc"),
                           os.path.join(ROOT, "tests", "cpp", "block_context_test.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
