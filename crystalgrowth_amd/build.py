"""Build the gfx950 library and the host-side C++ programs in-tree.

    python -m crystalgrowth_amd.build          # compile everything
    python -m crystalgrowth_amd.build --check  # report what is built

Outputs (next to their sources, git-ignored, shipped with the tree):
  crystalgrowth_amd/libkobayashi_gpu.so      C ABI of include/kobayashi_gpu.h
  crystalgrowth_amd/bin/kobayashi_headless   headless driver of the C++ Kobayashi class
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libkobayashi_gpu.so")
BIN = os.path.join(PKG, "bin")
HEADLESS = os.path.join(BIN, "kobayashi_headless")

ARCH = "gfx950"
HIP_SOURCES = ["kg_step.hip", "kg_api.hip", "kg_batch.hip", "kg_adjoint.hip"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 (v_pk_*_f32) stays off: hand-written HIP must not rely on it
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the scheduler's AMDGPU register-pressure trackers: fewer scalar
    # registers spilled in the step kernels, and the dense step 0.3-1.2 %
    # faster (four A/B pairs on MI355X); the arithmetic is unchanged
    "-mllvm", "-amdgpu-use-amdgpu-trackers=1",
    "-Wall", "-Wno-unused-function",
]


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build crystalgrowth_amd)")


def _run(cmd: list[str], quiet: bool) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # the host pass of hipcc does not know the device-only feature flag
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + "\n".join(lines))
    if lines and not quiet:
        print("\n".join(lines))


def build_library(quiet: bool = True) -> str:
    """hipcc the kernels and the C ABI into one shared library."""
    objs = []
    tmpdir = os.path.join(PKG, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    hipcc = _hipcc()
    for src in HIP_SOURCES:
        obj = os.path.join(tmpdir, src.replace(".hip", ".o"))
        _run([hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj], quiet)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], quiet)
    os.replace(tmp, LIB)
    return LIB


def build_headless(quiet: bool = True) -> str:
    """g++ the host C++ Kobayashi class and its headless driver against the library."""
    os.makedirs(BIN, exist_ok=True)
    srcs = [os.path.join(CSRC, "Kobayashi.cpp"), os.path.join(CSRC, "kobayashi_headless.cpp")]
    tmp = HEADLESS + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", CSRC, *srcs,
          "-o", tmp, "-L", PKG, "-lkobayashi_gpu", "-Wl,-rpath,$ORIGIN/.."], quiet)
    os.replace(tmp, HEADLESS)
    return HEADLESS


def build_all(quiet: bool = True) -> None:
    build_library(quiet)
    build_headless(quiet)


def is_built() -> bool:
    return os.path.exists(LIB)


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(LIB, "built" if os.path.exists(LIB) else "missing")
        print(HEADLESS, "built" if os.path.exists(HEADLESS) else "missing")
    else:
        build_all(quiet=False)
        print("built", LIB)
