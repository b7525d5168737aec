"""Build the native libraries in-tree.

  lib/libocl_hip.so        HIP kernels for gfx950 + the C ABI (include/ocl_hip.h)
  lib/libocl.so            C++ drop-in classes (ocl::BatchDropCutter, ...)
  lib/_ocl*.so             Python binding of the C++ classes (pybind11)
  lib/libocl_selcheck.so   host build of the selection arithmetic (CPU tests)

The oracle (oracle/build/libocl_oracle.so) is built by oracle/Makefile; it is
test infrastructure and nothing here links it.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OCL_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=off: the exact FP64 path must not fuse multiply-adds, so its
# results match the reference's x86-64 build; the selection kernels use
# explicit fma().  Packed FP32 instructions are disabled on purpose.
HIP_FLAGS = ["-O3", "-std=c++17", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fPIC"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    return r


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libocl_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, os.path.join(CSRC, "ocl_hip.hip")])
    return out


def build_selcheck():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libocl_selcheck.so")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", out,
          os.path.join(CSRC, "selcheck.cpp")])
    return out


def build_host():
    """C++ drop-in classes over the C ABI."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libocl.so")
    if not os.path.isdir(HOST):
        return None
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cpp")]
    if not srcs:
        return None
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-I", HOST, "-o", out, *srcs,
          "-L", LIB, "-locl_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_pybind():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HOST, "pybind", "ocl_module.cpp")
    if not os.path.exists(src):
        return None
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(LIB, "_ocl" + suffix)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE, "-I", HOST,
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"], "-o", out, src,
          "-L", LIB, "-locl", "-locl_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(verbose=False):
    steps = [("oracle", build_oracle), ("hip", build_hip), ("selcheck", build_selcheck),
             ("host", build_host), ("pybind", build_pybind)]
    for name, fn in steps:
        out = fn()
        if verbose:
            print(f"[build] {name}: {out or 'ok'}", flush=True)


if __name__ == "__main__":
    build_all(verbose=True)
