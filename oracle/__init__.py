"""CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built from ``oracle/trv_oracle.cpp`` by
``oracle/Makefile``), a from-scratch FP64 restatement of the reference's
three-point path.  Only the test suite, the smoke check and the benchmark's
CPU-baseline leg may import this package, and only to check or time the
reference algorithm: it is never the thing measured as the engine, and the
engine never calls it.
"""

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'liboracle.so')
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from triumvirate_amd import _abi
    lib = C.CDLL(LIB_PATH)
    _abi.declare_common(lib, 'trvo_')
    lib.trvo_set_schedule.argtypes = [C.c_int]
    lib.trvo_set_threads.argtypes = [C.c_int]
    lib.trvo_get_threads.restype = C.c_int
    dp = C.POINTER(C.c_double)
    lib.trvo_ylm.argtypes = [C.c_int, C.c_int, dp, dp]
    lib.trvo_wigner_3j.argtypes = [C.c_int] * 6
    lib.trvo_wigner_3j.restype = C.c_double
    lib.trvo_sph_bessel.argtypes = [C.c_int, C.c_double]
    lib.trvo_sph_bessel.restype = C.c_double
    lib.trvo_bessel_spline.argtypes = [C.c_int, dp, dp, C.c_longlong]
    lib.trvo_fft3d.argtypes = [dp, C.POINTER(C.c_int), C.c_int]
    lib.trvo_assign.argtypes = [C.POINTER(_abi.Catalogue), dp,
                                C.POINTER(_abi.Params), C.c_int, dp]
    lib.trvo_assign.restype = C.c_int
    lib.trvo_fourier_transform.argtypes = [C.POINTER(_abi.Params), dp, dp]
    lib.trvo_fourier_transform.restype = C.c_int
    lib.trvo_window.argtypes = [C.POINTER(_abi.Params), dp]
    lib.trvo_window.restype = C.c_int
    lib.trvo_alias.argtypes = [C.POINTER(_abi.Params), dp]
    lib.trvo_alias.restype = C.c_int
    lib.trvo_bench_pairs.argtypes = [C.POINTER(_abi.Params),
                                     C.POINTER(_abi.Binning), C.c_int, dp]
    lib.trvo_bench_pairs.restype = C.c_int
    _lib = lib
    return lib


def backend(schedule=0, threads=None):
    """Host-driver backend running the oracle (schedule 0: the reference's
    loop structure; 1: deduplicated shells)."""
    from triumvirate_amd.threept import CLibBackend
    lib = load()
    lib.trvo_set_schedule(int(schedule))
    if threads:
        lib.trvo_set_threads(int(threads))
    return CLibBackend(lib, 'trvo_')
