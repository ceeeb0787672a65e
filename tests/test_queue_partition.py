"""The persistent kernels' work-queue partition (fn2::queuePart) on the CPU:
tests/cpp/queue_partition.cpp simulates every wave of a launch, in random
interleavings, for grids of 1 to 70 waves and larger ones, and checks that
each item runs once and that every wave leaves."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_queue_partition_covers_every_item_once(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ is not installed")
    exe = str(tmp_path / "queue_partition")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "fastnoise2_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "queue_partition.cpp"), "-o", exe], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
