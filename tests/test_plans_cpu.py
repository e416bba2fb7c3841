"""Launch plans of the tiled step, checked on the CPU: tests/native/plans.cpp
builds the host-only planner (lgca_amd/csrc/lgca_plan.cpp) with the host
compiler and checks the strip-mode wave order (a permutation; solid walks,
then forced strips, then the rest in contiguous runs per XCD) and the band
plan of multi-pass launches (contiguous bands of whole tiles covering the
rows once) over lattices in strip and segment mode, whole and short last
words, steps per pass 1 to 8 and tile heights 2 to 1000."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def test_launch_plans(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.exists(os.path.join(ROCM, "include", "hip", "hip_runtime.h")):
        pytest.skip("g++ or the HIP headers are not available")
    exe = tmp_path / "plans"
    csrc = os.path.join(ROOT, "lgca_amd", "csrc")
    r = subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROCM, "include"), "-I", csrc,
                        os.path.join(ROOT, "tests", "native", "plans.cpp"), os.path.join(csrc, "lgca_plan.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("PLANS PASSED"), r.stdout[-4000:] + r.stderr
