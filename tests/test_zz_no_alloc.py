"""GPU: the no-allocation entry point against the allocating one.

gcomcot_detmath_eval_noalloc_ stages its arguments and results in the
library's static scratch array instead of device memory it allocates.
Two calls in a row, the second with fewer arguments than the first, must
each give exactly what gcomcot_detmath_eval_ gives on the same arguments:
the second call finds the first call's values still in the scratch array
beyond its own n, and must neither read nor return them.
"""
import ctypes

import numpy as np
import pytest

from gcomcot._lib import ip, last_error, lib

pytestmark = pytest.mark.gpu

_F = ctypes.POINTER(ctypes.c_float)
MAX_ARGS = 1 << 20


def _eval(fn, x):
    n = ctypes.c_longlong(x.size)
    pw = np.full_like(x, np.nan)
    ex = np.full_like(x, np.nan)
    rc = fn(ctypes.byref(n), x.ctypes.data_as(_F), pw.ctypes.data_as(_F), ex.ctypes.data_as(_F))
    assert rc == 0, f"return code {rc}"
    return pw, ex


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_no_alloc_equals_the_allocating_entry_point():
    lib().gcomcot_set_error_mode_(ip(0))
    rng = np.random.default_rng(7)
    first = np.exp(rng.uniform(np.log(1e-5), np.log(3e4), MAX_ARGS)).astype(np.float32)
    second = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), 1000 + 13)).astype(np.float32)
    for k, x in enumerate((first, second)):
        print(f"[gpu] no-alloc call {k + 1}: n = {x.size}", flush=True)
        pa, ea = _eval(lib().gcomcot_detmath_eval_, x)
        pn, en = _eval(lib().gcomcot_detmath_eval_noalloc_, x)
        assert np.array_equal(_bits(pa), _bits(pn)), f"call {k + 1}: powers differ"
        assert np.array_equal(_bits(ea), _bits(en)), f"call {k + 1}: exponentials differ"
        assert not np.isnan(pn).any() and not np.isnan(en).any()
    # beyond the maximum: refused before anything runs
    n = ctypes.c_longlong(MAX_ARGS + 1)
    one = np.ones(1, np.float32)
    assert lib().gcomcot_detmath_eval_noalloc_(ctypes.byref(n), one.ctypes.data_as(_F),
                                               one.ctypes.data_as(_F),
                                               one.ctypes.data_as(_F)) == 37
    assert last_error()[0] == 37  # the refusal is reported once
    assert one[0] == 1.0
