"""Sponge layer and wall boundary: library helpers and oracle against the reference.

SPONGE_COEF and BC_WALL run once on the host at set-up (comcot.f90:220-221);
the library's gcomcot_sponge_coef_ / gcomcot_bc_wall_ and the oracle's
restatement must reproduce the flang-built reference's arrays bit for bit.
SPONGE_LAYER acts after every momentum pass (comcot.f90:403-406); the
oracle's step with it must equal the reference's step with it bit for bit
on Cartesian layers (on spherical layers the reference's coefficients are
all 0, which both sides must also show).
"""
import ctypes

import numpy as np
import pytest

from gcomcot import lib
from gcomcot import scenarios as S
from helpers import BEACH, diff_report
from oracle import sponge as SP
from oracle.bind import Oracle, Reference

needs_ref = pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i(v):
    return ctypes.byref(ctypes.c_int(v))


def _library_coef(L, x, y):
    h = np.ascontiguousarray(L.h, np.float32)
    cx = np.zeros(L.nx, np.float32)
    cy = np.zeros(L.ny, np.float32)
    xa = np.ascontiguousarray(x, np.float32)
    ya = np.ascontiguousarray(y, np.float32)
    lib().gcomcot_sponge_coef_(_i(L.laycord), _i(L.nx), _i(L.ny), _p(h), _p(xa), _p(ya), _p(cx),
                               _p(cy))
    return cx, cy


CASES = {
    "cart-nonlinear": dict(laycord=1, laygov=1, nx=96, ny=80, **BEACH),
    "cart-linear": dict(laycord=1, laygov=0, nx=96, ny=80, cx=0.45),
    "sph-linear": dict(laycord=0, laygov=0, nx=96, ny=80, cx=0.45),
}


@needs_ref
@pytest.mark.parametrize("case", sorted(CASES))
def test_sponge_coefficients_match_the_reference(case):
    sc = S.custom(**CASES[case])
    L = sc.layers[0]
    x, y = S.layer_coords(L)
    ref = Reference(sc)
    ref.set_coords(1, x, y)
    rcx, rcy = ref.sponge(0)
    lcx, lcy = _library_coef(L, x, y)
    ocx, ocy = SP.sponge_coef(L.laycord, L.h, x, y)
    want_x = np.broadcast_to(lcx[None, :], rcx.shape)
    want_y = np.broadcast_to(lcy[:, None], rcy.shape)
    assert diff_report(rcx, want_x)[0] == 0 and diff_report(rcy, want_y)[0] == 0
    assert diff_report(ocx, lcx)[0] == 0 and diff_report(ocy, lcy)[0] == 0
    if L.laycord == 1:
        assert (lcx > 0).sum() >= 4 and (lcy > 0).sum() >= 4 and (lcx == 0).any()
    else:
        assert not lcx.any() and not lcy.any()   # degrees against metres: all 0


@needs_ref
@pytest.mark.parametrize("case", ["cart-nonlinear", "cart-linear"])
def test_oracle_step_with_sponge_matches_the_reference(case):
    sc = S.custom(**CASES[case])
    L = sc.layers[0]
    x, y = S.layer_coords(L)
    ref = Reference(sc)
    ref.set_coords(1, x, y)
    ref.sponge(1)
    orc = Oracle(sc, 32, pow_mode=0)
    orc.set_sponge(*SP.sponge_coef(L.laycord, L.h, x, y))
    plain = Oracle(sc, 32, pow_mode=0)
    n = 60
    ref.step(n)
    orc.step(n)
    plain.step(n)
    r, o, p = ref.state(1), orc.state(1), plain.state(1)
    for k in ("z1", "m1", "n1"):
        nb, mx = diff_report(o[k], r[k])
        assert nb == 0, f"{k}: {nb} cells differ (max {mx})"
    assert diff_report(p["m1"], o["m1"])[0] > 0, "the sponge must change the fluxes"


@needs_ref
def test_wall_matches_the_reference():
    sc = S.custom(laycord=1, laygov=1, nx=40, ny=32, **BEACH)
    L = sc.layers[0]
    ref = Reference(sc)
    ref.bc_wall()
    h = np.ascontiguousarray(L.h, np.float32).copy()
    lib().gcomcot_bc_wall_(_p(h), _i(L.nx), _i(L.ny))
    np.testing.assert_array_equal(h, ref.deform(1)[1])
    assert (h[:2] == -999).all() and (h[:, -2:] == -999).all() and (h[5:-5, 5:-5] != -999).all()
