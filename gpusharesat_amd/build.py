"""Builds libgpusharesat.so in-tree with hipcc for gfx950.

Usage: python -m gpusharesat_amd.build   (or gpusharesat_amd.build.build())

Every HIP source is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host pass warns that it
does not know the feature, which is harmless.
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
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libgpusharesat.so")

ARCH = os.environ.get("GSS_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "kernels.hip",
    "engine.hip",
    "residency.hip",
    "ceilings.hip",
    "host_state.cpp",
    "sharer.cpp",
    "sharer_capi.cpp",
    "serve.cpp",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    # keep real warnings visible, drop the known harmless host-pass notice
    for line in (r.stdout + r.stderr).splitlines():
        if "packed-fp32-ops" in line or not line.strip():
            continue
        if "warning" in line or "error" in line:
            sys.stderr.write(line + "\n")


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", obj]
        else:
            # host-only C++ that calls the HIP runtime API
            rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
            cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include", "-x", "c++",
                   "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
