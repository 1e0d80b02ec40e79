"""One HIP runtime per process.

PyTorch-ROCm wheels ship their own libamdhip64.so and load it by that unversioned name, so a
library that pulled in /opt/rocm's libamdhip64.so.7 first would leave the process with two
HIP runtimes, and the second one finds no GPU.  Before any native part of this package is
loaded, bind_hip_runtime() loads the runtime PyTorch will use (when PyTorch is installed);
libnkpr.so and libnkgen.so then resolve their libamdhip64.so.7 dependency to that same copy.
Without PyTorch, the system ROCm runtime is used.
"""
import ctypes
import importlib.util
import os

_bound = None


def bind_hip_runtime():
    global _bound
    if _bound is not None:
        return _bound
    path = None
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is not None and spec.origin:
        cand = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
        if os.path.exists(cand):
            ctypes.CDLL(cand, mode=ctypes.RTLD_GLOBAL)
            path = cand
    _bound = path or "system"
    return _bound
