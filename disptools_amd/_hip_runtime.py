"""One HIP runtime per process.

PyTorch-ROCm ships its own HIP runtime under ``torch/lib`` with the same
soname as ``/opt/rocm``'s.  Whichever is mapped first is the one every later
library binds to: if this package's library bound ``/opt/rocm``'s and torch
were imported afterwards, torch would find that runtime in place and see no
device.  ``preload()`` therefore maps torch's runtime first when torch is
installed -- by path, without importing torch -- so the library and a torch
imported later share it (device pointers of torch tensors are then valid in
the library).  Without torch it does nothing and the library binds
``/opt/rocm``'s runtime.
"""

import ctypes
import importlib.util
import os
import sys

_done = False


def preload():
    global _done
    if _done:
        return
    _done = True
    if "torch" in sys.modules:
        return   # torch is loaded: its runtime is in place
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.submodule_search_locations:
        return
    lib_dir = os.path.join(list(spec.submodule_search_locations)[0], "lib")
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        path = os.path.join(lib_dir, name)
        if os.path.exists(path):
            ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
