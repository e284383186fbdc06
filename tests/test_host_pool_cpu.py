"""The drop-in calls' host thread pool (csrc/wenox_hostpool.h): a small C++
driver runs jobs of every size and thread count, synchronously and while the
calling thread works, and checks that each row range is visited once; built
with ThreadSanitizer where g++ supports it, so a data race fails the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "weno-cfd-solver_amd", "csrc")
SRC = os.path.join(HERE, "cpp", "host_pool_test.cpp")


def _build(tmp_path, flags):
    exe = str(tmp_path / "host_pool_test")
    cmd = ["g++", "-std=c++17", "-O1", "-g", *flags, "-I", CSRC, SRC, "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return exe if r.returncode == 0 else None


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_pool_jobs(tmp_path):
    exe = _build(tmp_path, ["-fsanitize=thread"]) or _build(tmp_path, [])
    assert exe, "host_pool_test.cpp did not build"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr
