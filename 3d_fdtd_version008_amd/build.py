"""Builds the gfx950 library lib3dfdtd_hip.so in-tree.

The host setup (fdtd_setup.cpp) and the auto-mesher (fdtd_automesh.cpp) are
compiled by g++ with -ffp-contract=off so their arrays keep the reference's
rounding; the device library
(fdtd_device.hip) by hipcc for gfx950 with packed-FP32 instruction selection
disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib3dfdtd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    tmp = os.path.join(HERE, "_build")
    os.makedirs(tmp, exist_ok=True)
    setup_o = os.path.join(tmp, "fdtd_setup.o")
    dev_o = os.path.join(tmp, "fdtd_device.o")
    mesh_o = os.path.join(tmp, "fdtd_automesh.o")
    for src, obj in (("fdtd_setup.cpp", setup_o), ("fdtd_automesh.cpp", mesh_o)):
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-c",
              os.path.join(CSRC, src), "-o", obj])
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, "fdtd_device.hip"), "-o", dev_o])
    tmp_lib = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", dev_o, setup_o, mesh_o, "-o", tmp_lib, "-pthread"])
    os.replace(tmp_lib, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
