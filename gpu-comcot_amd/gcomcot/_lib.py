"""ctypes binding of libgcomcot_hip.so (the C ABI of include/gcomcot_hip.h).

The library is built in-tree by gpu-comcot_amd/Makefile.  Loading fails
loudly when it is missing: there is no CPU fallback for the step.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcomcot_hip.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_float_p = ctypes.POINTER(ctypes.c_float)

_lib = None


class GcomcotError(RuntimeError):
    pass


def lib():
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GcomcotError(
                f"{LIB_PATH} is missing: build it first (make -C gpu-comcot_amd or "
                "__graft_entry__.build()); the step has no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        L.gcomcot_init_layer_ex_.restype = ctypes.c_int
        L.gcomcot_run_steps_.restype = ctypes.c_int
        L.gcomcot_time_steps_.restype = ctypes.c_int
        L.gcomcot_last_error_.restype = ctypes.c_int
        L.gcomcot_layer_dims_.restype = ctypes.c_int
        L.gcomcot_cr_dt_.restype = ctypes.c_float
        L.gcomcot_set_tile_variant_.restype = ctypes.c_int
        L.gcomcot_tile_variants_.restype = ctypes.c_int
        L.gcomcot_hbm_copy_.restype = ctypes.c_int
        L.gcomcot_set_temporal_blocking_.restype = ctypes.c_int
        L.gcomcot_autotune_.restype = ctypes.c_int
        for name in ("gcomcot_set_layer_coords_", "gcomcot_gauges_set_", "gcomcot_gauges_read_"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def ip(v):
    """Fortran-style by-reference int argument."""
    return ctypes.byref(ctypes.c_int(int(v)))


def fp(v):
    """Fortran-style by-reference float argument."""
    return ctypes.byref(ctypes.c_float(float(v)))


def arr(a, dtype=np.float32):
    """Pointer to a contiguous array of the given dtype (no copy if already so)."""
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def last_error():
    buf = ctypes.create_string_buffer(512)
    code = lib().gcomcot_last_error_(buf, ip(512))
    return code, buf.value.decode(errors="replace")


def check(code, what):
    if code != 0:
        c, msg = last_error()
        raise GcomcotError(f"{what} failed (code {code}): {msg}")
