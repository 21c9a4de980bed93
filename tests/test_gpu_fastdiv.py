"""GPU: the short divisions and square root of the nonlinear momentum kernels.

div12 and sqrt_inrange over every float of their ranges, and div_inrange
over 2^30 pseudo-random operand pairs spread across the whole range where
it claims to equal IEEE division (gpu-comcot_amd/csrc/fastdiv.h), each
against the device's IEEE operation.
"""
import ctypes

import numpy as np
import pytest

from gcomcot import lib
from gcomcot._lib import check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 20260])
def test_short_divisions_and_sqrt_equal_ieee(seed):
    print(f"[gpu] division check seed {seed}", flush=True)
    pairs = ctypes.c_longlong(1 << 30)
    sd = ctypes.c_ulonglong(seed)
    b12, bin_, bsq = (ctypes.c_longlong(-1) for _ in range(3))
    tested = ctypes.c_longlong(0)
    ex, ea, eb, es = (ctypes.c_uint(0) for _ in range(4))
    check(lib().gcomcot_division_check_(ctypes.byref(pairs), ctypes.byref(sd), ctypes.byref(b12),
                                        ctypes.byref(bin_), ctypes.byref(bsq), ctypes.byref(tested),
                                        ctypes.byref(ex), ctypes.byref(ea), ctypes.byref(eb),
                                        ctypes.byref(es)),
          "division check")
    f = lambda u: float(np.array([u], np.uint32).view(np.float32)[0])
    assert tested.value >= 1 << 30
    assert b12.value == 0, f"div12: {b12.value} mismatches, e.g. x = {f(ex.value)!r}"
    assert bsq.value == 0, f"sqrt_inrange: {bsq.value} mismatches, e.g. x = {f(es.value)!r}"
    assert bin_.value == 0, (f"div_inrange: {bin_.value} mismatches, e.g. "
                             f"{f(ea.value)!r} / {f(eb.value)!r}")
