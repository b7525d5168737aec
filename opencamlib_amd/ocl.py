"""The OpenCAMLib Python interface on the MI355X kernels.

Re-exports the pybind11 module built from opencamlib_amd/host/pybind
(lib/_ocl*.so): Point, CLPoint, CCPoint, CCType, Triangle, STLSurf,
STLReader, the cutters, BatchDropCutter, PointDropCutter, BatchPushCutter,
Fiber, Interval, Line, Arc, Path, PathDropCutter, AdaptivePathDropCutter.
Raises ImportError when the native module has not been built.
"""
from __future__ import annotations

import glob
import importlib.machinery
import importlib.util
import os

from ._runtime import preload_hip_runtime

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


def _load():
    found = sorted(glob.glob(os.path.join(_LIB, "_ocl*.so")))
    if not found:
        raise ImportError(f"native module _ocl missing in {_LIB}; run __graft_entry__.build()")
    preload_hip_runtime()
    loader = importlib.machinery.ExtensionFileLoader("_ocl", found[0])
    spec = importlib.util.spec_from_file_location("_ocl", found[0], loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


_mod = _load()
globals().update({k: getattr(_mod, k) for k in dir(_mod) if not k.startswith("__")})
Error = _mod.Error
