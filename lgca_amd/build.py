"""Builds liblgca_hip.so (gfx950) in-tree with hipcc.

    python -m lgca_amd.build            # or lgca_amd.build.build()

Objects go to lgca_amd/_build/, the shared library to lgca_amd/liblgca_hip.so.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblgca_hip.so")

SOURCES = ["lgca_kernels.hip", "lgca_tile.hip", "lgca_capi.hip", "lgca_batch.hip", "lgca_plan.cpp"]
# lgca_tile_inst.hip is compiled once per (model kind, steps per pass) pair of
# the single-lattice kernels and once per steps per pass of the batch kernel
# (kind -1), which lets the template-heavy tiled kernels compile in parallel
TILE_INST = [(kind, k) for k in (8, 4, 2, 1) for kind in (-1, 0, 1, 2)]
ARCH = os.environ.get("LGCA_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 VALU ops stay disabled in every gfx950 build of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def _hipcc() -> str:
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _compile(job) -> str:
    src, defines, name = job
    obj = os.path.join(BUILD, name + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} {' '.join(defines)}:\n{r.stdout}\n{r.stderr}")
    return obj


def _jobs():
    jobs = [(f"lgca_tile_inst.hip", [f"-DLGCA_INST_KIND={kind}", f"-DLGCA_INST_K={k}"],
             f"lgca_tile_{'batch' if kind < 0 else f'm{kind}'}_k{k}") for kind, k in TILE_INST]
    jobs += [(src, [], os.path.splitext(src)[0]) for src in SOURCES]
    return jobs


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = _jobs()
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
