"""Device pow/exp of the nonlinear momentum against the oracle's host build.

The momentum kernels raise depths to 2.3333333 for Manning friction
(moment.f90:459-460, 945-946) and take EXP(-REF) for wave breaking
(moment.f90:1323).  Both are computed in double from one source
(csrc/detmath.h) on host and device; this test checks, bit for bit, that
the device code gives what the oracle's host build gives, over depths
spread log-uniformly across the range the callers pass (x >= 1e-5), over
the breaking arguments, and over raw float bit patterns for the
exponential (which takes every float).
"""
import ctypes

import numpy as np
import pytest

from gcomcot._lib import ip, lib
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

MAN_EXP = np.float32(2.3333333)


@pytest.fixture(autouse=True)
def _errors_return():
    lib().gcomcot_set_error_mode_(ip(0))     # errors return codes instead of exiting


def _device(x):
    n = ctypes.c_longlong(x.size)
    pw = np.empty_like(x)
    ex = np.empty_like(x)
    f = ctypes.POINTER(ctypes.c_float)
    rc = lib().gcomcot_detmath_eval_(ctypes.byref(n), x.ctypes.data_as(f), pw.ctypes.data_as(f),
                                     ex.ctypes.data_as(f))
    assert rc == 0
    return pw, ex


def _host(x):
    L = Oracle.lib()
    f = ctypes.POINTER(ctypes.c_float)
    L.orc_detmath_eval.restype = None
    L.orc_detmath_eval.argtypes = [ctypes.c_longlong, f, ctypes.c_float, f, f]
    pw = np.empty_like(x)
    ex = np.empty_like(x)
    L.orc_detmath_eval(x.size, x.ctypes.data_as(f), float(MAN_EXP), pw.ctypes.data_as(f),
                       ex.ctypes.data_as(f))
    return pw, ex


def _same(a, b):
    return np.ascontiguousarray(a).view(np.uint32) == np.ascontiguousarray(b).view(np.uint32)


def test_device_pow_and_exp_equal_the_host_build():
    rng = np.random.default_rng(20261202)
    depths = np.exp(rng.uniform(np.log(1e-5), np.log(3e4), 1 << 21)).astype(np.float32)
    edges = np.array([1e-5, 1e-3, 1.0, 2.0, 1e4, 3e16, 3e38, np.inf, np.nan], np.float32)
    ref = -rng.uniform(0.0, 40.0, 1 << 19).astype(np.float32)           # EXP(-REF), REF >= 0
    bits = rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x = np.concatenate([depths, edges, ref, bits]).astype(np.float32)
    print(f"[gpu] detmath eval on {x.size} arguments", flush=True)
    dpw, dex = _device(x)
    hpw, hex_ = _host(x)
    on_domain = ~(x < 1e-5)                 # the power's callers pass x >= 1e-5 (NaN included)
    bad_pow = np.flatnonzero(on_domain & ~_same(dpw, hpw))
    bad_exp = np.flatnonzero(~_same(dex, hex_))
    assert bad_pow.size == 0, f"pow differs at x = {x[bad_pow[:5]]}: {dpw[bad_pow[:5]]} vs {hpw[bad_pow[:5]]}"
    assert bad_exp.size == 0, f"exp differs at x = {x[bad_exp[:5]]}: {dex[bad_exp[:5]]} vs {hex_[bad_exp[:5]]}"
    # the values are the functions themselves, not a shared wrong answer
    d = depths[:4096].astype(np.float64)
    rel = np.abs(dpw[:4096] / d ** np.float64(MAN_EXP) - 1.0)
    assert rel.max() < 1.2e-7
    r = ref[:4096].astype(np.float64)
    k = slice(depths.size + edges.size, depths.size + edges.size + 4096)
    assert np.abs(dex[k] / np.exp(r) - 1.0).max() < 1.2e-7

