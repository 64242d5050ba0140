"""ctypes binding of the native engine library (_lib/libhfx.so, built by build.py).

The library is the product: there is no Python or CPU fallback.  Importing
this module on a machine where the library is missing raises ImportError with
the build command; calling into it without a visible GPU returns the HIP
error as an HfxError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# HFX_LIBRARY selects another build of the same library (the CPU sanitizer
# builds of tests/sanitize); unset, the in-tree product build is loaded.
LIB_PATH = os.environ.get("HFX_LIBRARY") or os.path.join(HERE, "_lib", "libhfx.so")

HFX_MODE_STORED, HFX_MODE_DIRECT, HFX_MODE_AUTO = 0, 1, 2
HFX_PREC_FP64, HFX_PREC_MIXED = 0, 1


class HfxError(RuntimeError):
    """A non-zero status from the native library, with its message."""

    def __init__(self, code, msg):
        super().__init__(f"hfx error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library {LIB_PATH} is missing; build it with "
                          "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
    # torch ships its own libamdhip64.so.7; importing it first makes the engine
    # bind to that same HIP runtime instead of starting a second one in-process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(LIB_PATH)


lib = _load()

_vp = C.c_void_p
lib.hfx_last_error.restype = C.c_char_p
lib.hfx_device_count.restype = C.c_int


class FockOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("precision", C.c_int), ("threshold", C.c_double),
                ("mixed_split", C.c_double), ("hbm_budget", C.c_size_t), ("incremental", C.c_int),
                ("incremental_tau", C.c_double)]


def check(rc):
    if rc != 0:
        raise HfxError(rc, lib.hfx_last_error().decode())
    return rc


def ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
