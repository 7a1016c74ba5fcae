"""Builds libgproshan_ptp.so in-tree (gproshan_amd/libgproshan_ptp.so).

The HIP translation units are compiled by hipcc for gfx950 only; the host
C++ units by g++ with OpenMP.  No CMake, no JIT: the built library is what
tests, smoke() and bench.py load.

Flags of note:
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: hipcc emits no packed
    FP32 (v_pk_*_f32) instructions.
  * ``-mllvm -amdgpu-kernarg-preload-count=16``: a kernel's leading plain
    arguments (up to 16 dwords) arrive in scalar registers at wave launch, so
    the sweep kernels' first dependent load (the schedule slot) does not wait
    for the kernel-argument segment.  Firmware without the feature runs the
    compiler's compatibility prologue, which loads them as before.
  * ``-ffp-contract=off`` on the host units: the host FP64 update rounds like
    the reference's default x86-64 build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgproshan_ptp.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["ptp_gpu.hip", "patches.hip", "rings.hip", "mlp.hip", "batch.hip", "fps.hip", "levels.hip",
               "noalloc.hip"]
CPP_SOURCES = ["che.cpp", "ptp_cpu.cpp", "capi.cpp", "mesh_io.cpp"]

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-mllvm", "-amdgpu-kernarg-preload-count=16",
               "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wextra",
             "-Wno-unused-parameter", f"-I{ROCM}/include"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    cxx = os.environ.get("CXX", "g++")
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([cxx, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([cxx, "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64", "-fopenmp",
          f"-Wl,-rpath,{ROCM}/lib", "-Wl,--no-undefined"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
