"""CPU: argument checking of the C ABI in status mode (no device needed).

Every entry point checks its arguments before it touches the device, so
these run without a GPU: each invalid call must return (or record) its
error code and a message, and leave nothing behind.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from gcomcot._lib import fp, ip, last_error, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(autouse=True)
def status_mode():
    lib().gcomcot_set_error_mode_(ip(0))
    last_error()
    yield
    last_error()


def init_ex(layerid=1, parent=-1, laycord=0, laygov=0, rel_size=1, rel_time=1,
            corners=(1, 1, 1, 1), fric_switch=1, fric_coef=0.0, dt=1.0, dx=1.0, dy=1.0,
            coef=True, nx=16, ny=12):
    h = np.zeros((ny, nx), np.float32)
    z = np.zeros((ny, nx), np.float32)
    c = np.zeros(10 * ny, np.float32)
    cn = np.array(corners, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    return lib().gcomcot_init_layer_ex_(ip(layerid), ip(parent), ip(1), ip(laycord), ip(laygov),
                                        ip(rel_size), ip(rel_time), p(cn), ip(fric_switch),
                                        fp(fric_coef), fp(dt), fp(dx), fp(dy),
                                        p(c) if coef else None, p(h), p(z), ip(nx), ip(ny))


@pytest.mark.parametrize("kw,code", [
    (dict(laycord=2), 10),            # coordinate system
    (dict(laygov=2), 11),             # dispersive governing equations
    (dict(laycord=0, laygov=5), 11),  # LAYGOV 5 is Cartesian only
    (dict(fric_switch=2), 12),        # variable Manning
    (dict(dt=0.0), 13),
    (dict(dx=-1.0), 13),
    (dict(coef=False), 14),           # spherical layer without its PARA rows
    (dict(layerid=0), 1),
    (dict(layerid=20), 1),
    (dict(nx=2), 2),
    (dict(parent=3), 3),              # layer 1 must be the root
    (dict(layerid=2, parent=1, laycord=1, rel_size=2, nx=9, ny=9), 4),  # parent not there
])
def test_init_layer_rejects(kw, code):
    assert init_ex(**kw) == code
    c, msg = last_error()
    assert c == code and msg


def test_calls_on_missing_layers_report():
    L = lib()
    L.mass_launch_(None, None, None, ip(7))
    assert last_error()[0] == 15
    L.momt_launch_(None, None, None, ip(7))
    assert last_error()[0] == 16
    L.maxamp_launch_(ip(7))
    assert last_error()[0] == 17
    L.cuda_update_layer_(ip(7))
    assert last_error()[0] == 18
    assert L.gcomcot_run_steps_(ip(-1), ip(1)) == 23
    assert L.gcomcot_time_steps_(ip(0), None, None) == 23
    assert L.gcomcot_set_tile_variant_(ip(999)) == 1


def test_batch_argument_checks():
    L = lib()
    ll = lambda v: ctypes.byref(ctypes.c_longlong(v))
    sid = ctypes.c_int(0)
    assert L.gcomcot_batch_new_scenario_(ll(-5), ctypes.byref(sid)) == 30
    assert L.gcomcot_batch_new_scenario_(ll(10), ctypes.byref(sid)) == 0
    h = np.zeros((9, 9), np.float32)
    cn = np.array([4, 7, 4, 7], np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    bad = ctypes.c_int(sid.value + 100)
    rc = L.gcomcot_batch_add_layer_(ctypes.byref(bad), ip(1), ip(-1), ip(1), ip(0), ip(1), ip(1),
                                    p(cn), ip(1), fp(0), fp(1), fp(1), fp(1), None, p(h), p(h),
                                    ip(9), ip(9))
    assert rc == 31
    rc = L.gcomcot_batch_add_layer_(ctypes.byref(sid), ip(3), ip(2), ip(1), ip(0), ip(2), ip(1),
                                    p(cn), ip(1), fp(0), fp(1), fp(1), fp(1), None, p(h), p(h),
                                    ip(9), ip(9))
    assert rc == 32  # one level of nesting below the root
    rc = L.gcomcot_batch_add_layer_(ctypes.byref(sid), ip(1), ip(-1), ip(1), ip(2), ip(1), ip(1),
                                    p(cn), ip(1), fp(0), fp(1), fp(1), fp(1), None, p(h), p(h),
                                    ip(9), ip(9))
    assert rc == 11
    assert L.gcomcot_batch_set_tuning_(ip(1), ip(6)) == 34
    assert L.gcomcot_batch_remove_(ctypes.byref(bad)) == 31


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_detmath_eval_refuses_bad_arguments():
    n = ctypes.c_longlong(-1)
    assert lib().gcomcot_detmath_eval_(ctypes.byref(n), None, None, None) == 36
    assert last_error()[0] == 36
    n = ctypes.c_longlong(4)                       # a count without arrays
    assert lib().gcomcot_detmath_eval_(ctypes.byref(n), None, None, None) == 36
    assert last_error()[0] == 36
    n = ctypes.c_longlong(0)                       # nothing to do: no device needed
    assert lib().gcomcot_detmath_eval_(ctypes.byref(n), None, None, None) == 0


def test_detmath_eval_noalloc_refuses_bad_arguments():
    f = lib().gcomcot_detmath_eval_noalloc_
    n = ctypes.c_longlong(-1)
    assert f(ctypes.byref(n), None, None, None) == 36
    n = ctypes.c_longlong(4)                       # a count without arrays
    assert f(ctypes.byref(n), None, None, None) == 36
    one = np.ones(1, np.float32)
    p = one.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    n = ctypes.c_longlong((1 << 20) + 1)           # beyond the scratch array
    assert f(ctypes.byref(n), p, p, p) == 37
    assert last_error()[0] == 37
    n = ctypes.c_longlong(0)                       # nothing to do: no device needed
    assert f(ctypes.byref(n), None, None, None) == 0

