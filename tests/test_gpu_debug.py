"""GPU: the reference's differential-debug entry points.

Built with DEBUG_CORE, the reference's Fortran computes JNQ/NEWQ and JNZ
on the host, then calls edgeinterp_dbglaunch_ / jnz_dbglaunch_
(all_grids.f90:417, 490, 650), which run the device's own routine on the
state SET_ALL copied over and print every value that differs from the
host's by 1e-4 or more (GPUall_grids.cu:299-394, 626-643).  Here the host
side is the oracle's JNQ, NEWQ and JNZ on the device's own state: the
checks must print no difference, and exactly the one planted in the host
arrays when there is one.
"""
import ctypes

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot._lib import ip, lib
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_P)


def _nest():
    return S.custom(nx=72, ny=64, laygov=0, cx=0.4,
                    child=dict(laycord=0, laygov=0, footprint=(20, 36), rel_size=3, rel_time=3,
                               corner=(8, 14)))


def _host_from_device(sc, dev):
    orc = Oracle(sc, 32, pow_mode=1)
    for L in sc.layers:
        st = dev.state(L.id)
        st["dz1"] = np.zeros_like(st["z1"])
        st["dz2"] = np.zeros_like(st["z1"])
        orc.set_state(L.id, st)
    return orc


def _lines(text, tag):
    return [ln for ln in text.splitlines() if ln.startswith(tag)]


def test_edgeinterp_dbglaunch_agrees_with_the_host(capfd):
    lib().gcomcot_set_error_mode_(ip(0))
    sc = _nest()
    dev = DeviceRun(sc)
    dev.run(25)
    print("[gpu] edgeinterp_dbglaunch_ after 25 steps", flush=True)
    for t in (1, 2):
        orc = _host_from_device(sc, dev)
        orc.call("JNQ" if t == 1 else "NEWQ", 2, t)
        c = orc.state(2)
        mA = np.ascontiguousarray(c["m1"], np.float32)
        nA = np.ascontiguousarray(c["n1"], np.float32)
        xf, yf = orc.flux(1)
        xflux = np.ascontiguousarray(xf, np.float32).reshape(-1)
        yflux = np.ascontiguousarray(yf, np.float32).reshape(-1)
        p = dev.state(1)
        mO, nO = (np.ascontiguousarray(p[k], np.float32) for k in ("m1", "n1"))
        assert np.abs(mA[:, 0]).max() > 1e-3, "no flux across the child's left boundary"
        capfd.readouterr()
        lib().edgeinterp_dbglaunch_(_p(mO), _p(nO), ip(1), _p(mA), _p(nA), ip(2), ip(t),
                                    _p(yflux), _p(xflux))
        out = capfd.readouterr().out
        assert "==============lO:1, lA:2 =============" in out, out
        diffs = [ln for ln in out.splitlines() if ln.startswith("(")]
        assert not diffs, f"t={t}: {diffs[:5]}"
        # one planted difference on each checked edge is reported, once
        mA[7, 0] += 1.0
        nA[-1, 5] -= 1.0
        if t == 2:
            yflux[sc.layers[1].corners[2] + 1] += 1.0
        lib().edgeinterp_dbglaunch_(_p(mO), _p(nO), ip(1), _p(mA), _p(nA), ip(2), ip(t),
                                    _p(yflux), _p(xflux))
        out = capfd.readouterr().out
        assert len(_lines(out, "(left) i = 7,")) == 1 and len(_lines(out, "(top) i = 5,")) == 1
        assert len(_lines(out, "(left-flux)")) == (1 if t == 2 else 0)
        assert len([ln for ln in out.splitlines() if ln.startswith("(")]) == (3 if t == 2 else 2)
    dev.close()


def test_jnz_dbglaunch_agrees_with_the_host(capfd):
    lib().gcomcot_set_error_mode_(ip(0))
    sc = _nest()
    dev = DeviceRun(sc)
    dev.run(30)
    print("[gpu] jnz_dbglaunch_ after 30 steps", flush=True)
    orc = _host_from_device(sc, dev)
    orc.call("JNZ", 2, 0)
    z = np.ascontiguousarray(orc.state(1)["z2"], np.float32)
    IS, IE, JS, JE = sc.layers[1].corners
    before = dev.state(1)["z2"]
    assert not np.array_equal(before[JS - 1:JE, IS - 1:IE], z[JS - 1:JE, IS - 1:IE]), \
        "JNZ changes nothing here: the test would not see it run"
    capfd.readouterr()
    lib().jnz_dbglaunch_(_p(z), ip(1), ip(2))
    out = capfd.readouterr().out
    assert f"child corner {IS} {IE} {JS} {JE}" in out, out
    assert not _lines(out, "JNZ, layerid"), out[:500]
    z[JS + 2, IS + 3] += 0.5
    lib().jnz_dbglaunch_(_p(z), ip(1), ip(2))
    out = capfd.readouterr().out
    bad = _lines(out, "JNZ, layerid=2")
    assert len(bad) == 1 and f"(i,j)=({IS + 3},{JS + 2})" in bad[0], out[:500]
    dev.close()


def test_dbglaunch_refuses_a_layer_that_is_not_a_child():
    lib().gcomcot_set_error_mode_(ip(0))
    from gcomcot._lib import last_error
    sc = _nest()
    dev = DeviceRun(sc)
    z = np.zeros((sc.layers[0].ny, sc.layers[0].nx), np.float32)
    lib().jnz_dbglaunch_(_p(z), ip(2), ip(1))
    assert last_error()[0] == 38
    dev.close()
