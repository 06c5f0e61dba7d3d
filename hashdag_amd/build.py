"""Builds libhashdag_hip.so (HIP kernels for gfx950 + host C++) in-tree.

Every HIP translation unit is compiled for gfx950 only, without FP contraction (the
tracer and editors must match the reference's IEEE operation order bit for bit) and with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libhashdag_hip.so")

HIP_SOURCES = ["hd_context.hip", "hd_trace.hip", "hd_stream.hip", "hd_edit.hip", "hd_edit_api.hip", "hd_gc.hip",
               "hd_serve.hip", "hd_replay.hip", "hd_files.hip"]
CPP_SOURCES = ["hd_scene.cpp", "hd_factory.cpp", "hd_validate.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = COMMON + [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fgpu-flush-denormals-to-zero=false" if False else "-fno-gpu-flush-denormals-to-zero",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "unknown target feature" not in l and "packed-fp32" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    for src in CPP_SOURCES:
        o = os.path.join(OBJ, src.replace(".cpp", ".o"))
        objs.append(o)
        jobs.append([cxx, *COMMON, "-c", os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
