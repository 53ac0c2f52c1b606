"""Build libfwhip.so (HIP kernels for gfx950 + C host driver) in-tree.

hipcc compiles the .hip sources for gfx950, gcc the C driver; hipcc links
the shared library next to this file.  Run as a script or call build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libfwhip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # packed-FP32 instructions are kept out of all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
HIP_SOURCES = ["fw_kernels.hip", "fw_solve.hip", "fw_schur.hip", "fw_factor.hip", "fw_host.hip"]
C_SOURCES = ["hddm_driver.c", "fw_refapi.c"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        out = _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        if verbose and out.strip():
            print(out)
        objs.append(o)
    for src in C_SOURCES:
        o = os.path.join(OBJ, src.replace(".c", ".o"))
        _run(["gcc", "-O2", "-std=c99", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-c",
              os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    shutil.rmtree(OBJ, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
