"""The oracle against the flang-built reference on fresh seeded inputs.

Needs oracle/_ref/libcomcot_ref.so (built by __graft_entry__.build() where
the reference sources exist); skipped elsewhere.
"""
import numpy as np
import pytest

from gcomcot import scenarios as S
from helpers import bits, small_scenarios
from oracle.bind import Oracle, Reference

pytestmark = pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("name,sc,steps", small_scenarios(), ids=[c[0] for c in small_scenarios()])
def test_oracle_equals_reference(name, sc, steps, oracle_built):
    steps = min(steps, 80)
    o, r = Oracle(sc, 32, 0), Reference(sc)
    o.step(steps)
    r.step(steps)
    for L in sc.layers:
        a, b = o.state(L.id), r.state(L.id)
        for k in ("z1", "m1", "n1", "z2", "m2", "n2", "dz2"):
            assert np.array_equal(bits(a[k]), bits(b[k])), f"{name} layer {L.id} {k}"


def test_setup_products_equal_reference(oracle_built):
    """PQ_DEPTH and PARA planes of the restatement and of the library's
    per-row table against the Fortran's, bit for bit."""
    sc = S.custom(laycord=0, laygov=0, nx=40, ny=36, child=dict(
        laycord=0, laygov=1, footprint=(10, 12), rel_size=3, rel_time=4, corner=(6, 8)))
    o, r = Oracle(sc, 32, 0), Reference(sc)
    for L in sc.layers:
        so, sr = o.setup(L.id), r.setup(L.id)
        for k in ("hp", "hq"):
            assert np.array_equal(bits(so[k]), bits(sr[k]))
        assert np.array_equal(bits(so["r"]), bits(sr["r"]))
        assert np.array_equal(bits(so["r6"]), bits(sr["r6"]))
        rows = S.para_rows(L.laygov, L.dt, L.dx, L.y, L.dely).reshape(10, L.ny)
        # row table k -> PARA plane index: R0,R1,R11,F2|R2,R21,R22,R3,F4|R4,R5
        for k, p in ((0, 0), (1, 1), (2, 2), (4, 4), (5, 5), (6, 6), (8, 8)):
            assert np.array_equal(bits(np.repeat(rows[k][:, None], L.nx, 1)), bits(sr["r"][p]))
        if L.laygov == 1:
            assert np.array_equal(bits(np.repeat(rows[3][:, None], L.nx, 1)), bits(sr["r"][3]))
            assert np.array_equal(bits(np.repeat(rows[7][:, None], L.nx, 1)), bits(sr["r"][7]))
        else:
            r2 = (np.repeat(rows[3][:, None], L.nx, 1) * np.maximum(sr["hp"], np.float32(0)))
            assert np.array_equal(bits(r2.astype(np.float32)), bits(sr["r"][3]))
        assert np.array_equal(bits(rows[9]), bits(sr["r6"]))


def test_time_step_equals_reference(oracle_built):
    """CR_CHECK: the library's dt for the root matches the Fortran's."""
    from oracle.bind import Reference as R
    import ctypes
    for lg in (0, 1):
        sc = S.custom(laycord=0, laygov=lg, nx=40, ny=36)
        L = sc.layers[0]
        big = S.LayerSpec(**{**L.__dict__, "dt": 1.0e4})
        ref = R(S.Scenario("cr", [big], 0))
        dt = ctypes.c_float(0)
        rt = (ctypes.c_int * 4)()
        dts = (ctypes.c_float * 4)()
        R.lib().ref_cr_check(ctypes.byref(dt), rt, dts)
        mine = S.cr_dt(0, lg, L.h, L.dx, L.dy, float(L.y[0]), float(L.y[-1]), 1.0e4)
        assert np.float32(mine) == dt.value
