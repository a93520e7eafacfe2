"""Loader for the HIP engine library (``libtrvgpu.so``, built in-tree).

There is no fallback: if the library is missing or cannot be loaded, every
estimator call raises.  Build it with ``python -c "import __graft_entry__ as
g; g.build()"`` or ``make -C triumvirate_amd/csrc``.
"""

import ctypes as C
import os
import threading

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libtrvgpu.so')

_lock = threading.Lock()
_lib = None


class EngineError(RuntimeError):
    """Raised for device/runtime failures reported by the engine."""


def load():
    """Load (once) and return the ctypes handle of ``libtrvgpu.so``."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"HIP engine library not found at {LIB_PATH}; build it "
                "first (make -C triumvirate_amd/csrc).")
        lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        _abi.declare_common(lib, 'trvg_')
        lib.trvg_abi_version.restype = C.c_int
        lib.trvg_device_count.argtypes = [C.POINTER(C.c_int)]
        lib.trvg_device_count.restype = C.c_int
        _abi.declare_extensions(lib)
        if lib.trvg_abi_version() != 1:
            raise ImportError("libtrvgpu.so ABI version mismatch.")
        _lib = lib
        return lib
