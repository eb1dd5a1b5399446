"""CPU test of the parallel host copier of the host-memory paths
(minhashcuda_amd/csrc/host_copy.h): tests/native/host_copy_check.cpp copies
buffers of many shapes through it and compares them byte for byte, built
with g++ as is and under ThreadSanitizer (the workers' job hand-off)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "host_copy_check.cpp")
CSRC = os.path.join(ROOT, "minhashcuda_amd", "csrc")


def build_and_run(tmp_path, flags, rounds):
    exe = str(tmp_path / "host_copy_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I" + CSRC] + flags +
                   [SRC, "-o", exe], check=True)
    return subprocess.run([exe, str(rounds)], stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=600)


def test_copies_are_exact(tmp_path):
    res = build_and_run(tmp_path, [], 8)
    assert res.returncode == 0 and res.stdout.strip().endswith("ok"), res.stdout[-2000:]


def test_no_data_race_under_thread_sanitizer(tmp_path):
    res = build_and_run(tmp_path, ["-fsanitize=thread"], 1)
    if "unexpected memory mapping" in res.stdout or "FATAL: ThreadSanitizer" in res.stdout:
        pytest.skip("ThreadSanitizer cannot run in this environment: " + res.stdout[:200])
    assert "WARNING: ThreadSanitizer" not in res.stdout, res.stdout[-4000:]
    assert res.returncode == 0 and res.stdout.strip().endswith("ok"), res.stdout[-2000:]
