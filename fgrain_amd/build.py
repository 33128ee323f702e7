"""Build recipe of libfgrain_hip.so (gfx950) -- run by ``__graft_entry__.build()``.

hipcc compiles the kernels for gfx950 and the host C++ of the engine; the library is
written in-tree next to this file so it travels with the repository snapshot.
Floating-point contraction is off everywhere (-ffp-contract=off): the geometry
kernels must round exactly like the float reference evaluation they are tested
against.  Packed-FP32 instruction selection is disabled for device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfgrain_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(HERE, "..", "include")]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]
HOST_SOURCES = ["host.cpp", "engine.cpp", "adjoint_host.cpp", "film_host.cpp"]
DEVICE_SOURCES = ["kernels.hip", "batched.hip", "adjoint.hip", "synth.hip", "film.hip", "noalloc.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *COMMON, *DEVICE, "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
