"""Tile placement of the fused and reverse TV kernels, checked on the CPU:
tests/host/tiling_check.hip is compiled for the host (it calls only the
host side of the shared __host__ __device__ placement functions) and run."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_every_column_has_one_owner(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc is needed to build the host-side check")
    exe = str(tmp_path / "tiling_check")
    cmd = [hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "sci-algorithms_amd", "csrc"),
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", os.path.join(HERE, "host", "tiling_check.hip"),
           "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), run.stdout[-2000:]
