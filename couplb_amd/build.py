"""Build the MI355X library in-tree: couplb_amd/libcouplb_hip.so.

  python -m couplb_amd.build          (or __graft_entry__.build())

Sources: couplb_amd/csrc/*.hip (HIP for gfx950) and couplb_amd/host/*.cpp
(the fix's host C++), linked into one shared library whose C ABI is
declared in include/couplb_hip.h, include/couplb_mlp.h and
include/couplb_fix.h.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcouplb_hip.so")
ARCH = os.environ.get("COUPLB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled (-packed-fp32-ops): the fused step
# must not depend on them.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    obj_hip = os.path.join(bdir, "couplb_hip.o")
    obj_mlp = os.path.join(bdir, "mlp.o")
    obj_fix = os.path.join(bdir, "fix_couplb_core.o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, "csrc", "couplb_hip.hip"), "-o", obj_hip])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(HERE, "csrc", "mlp.hip"), "-o", obj_mlp])
    obj_vtk = os.path.join(bdir, "vtk_io.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c",
          os.path.join(HERE, "host", "fix_couplb_core.cpp"), "-o", obj_fix])
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c",
          os.path.join(HERE, "host", "vtk_io.cpp"), "-o", obj_vtk])
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, obj_hip, obj_mlp, obj_fix,
          obj_vtk])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
