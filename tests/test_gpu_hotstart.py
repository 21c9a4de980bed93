"""GPU: the reference's hot-start files written from and read into the device.

A nest runs on the device; gcomcot_hotstart_write_ puts down the five
snapshot files per layer (HOTSTART_DATA's names and '(10F12.5)' layout,
pinned against the flang build in test_hotstart_oracle.py); a fresh tree
loads them with gcomcot_hotstart_read_ and continues.  The continuation
must equal, bit for bit, the oracle started from the same rounded state
(level 2 equal to level 1, as the level copy leaves it), with arrival steps
counted on from START_STEP.  START_TYPE 1 loads the outermost layer only,
and a missing file zeroes its plane as HOTSTART_DATA does.
"""
import ctypes
import os

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot._lib import ip, last_error, lib
from helpers import BEACH, diff_report
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

KINDS = ("z1", "m1", "n1", "zmax1", "zmin1")


def _scenario():
    return S.custom(nx=72, ny=64, laygov=1, child=dict(laycord=1, laygov=1, footprint=(20, 36),
                                                       rel_size=3, rel_time=4, corner=(8, 14)),
                    **BEACH)


def _write(step, d):
    b = str(d).encode()
    assert lib().gcomcot_hotstart_write_(ip(step), b, ip(len(b))) == 0


def _read(start_type, step, d):
    b = str(d).encode()
    miss = ctypes.c_int(-1)
    rc = lib().gcomcot_hotstart_read_(ip(start_type), ip(step), b, ip(len(b)), ctypes.byref(miss))
    assert rc == 0, rc
    return miss.value


def _file(path, shape, first):
    a = np.zeros(shape, np.float32)
    b = str(path).encode()
    assert lib().gcomcot_hotstart_file_read_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                             ip(shape[1]), ip(shape[0]), ip(first)) == 0
    return a


def _same(a, b, tag):
    nb, mx = diff_report(a, b)
    assert nb == 0, f"{tag}: {nb} cells differ (max {mx})"


def test_hotstart_files_resume_to_the_oracle_bits(tmp_path):
    lib().gcomcot_set_error_mode_(ip(0))
    sc = _scenario()
    done, more = 31, 24
    start = done + 1
    print(f"[gpu] hot start: write after {done} steps, resume for {more}", flush=True)
    dev = DeviceRun(sc)
    dev.run(done)
    _write(start, tmp_path)
    want = sorted(f"{k}_{L.id:02d}_{start:06d}.dat" for L in sc.layers for k in KINDS)
    assert sorted(os.listdir(tmp_path)) == want
    before = {L.id: dev.state(L.id) for L in sc.layers}
    raw1 = dev.maxima(1)["zmax_raw"]
    dev.close()

    # the files hold the device state rounded to 5 decimals
    snap = {}
    for L in sc.layers:
        first = 1 if L.id == 1 else 2
        for k in KINDS:
            snap[(L.id, k)] = _file(tmp_path / f"{k}_{L.id:02d}_{start:06d}.dat",
                                    (L.ny, L.nx), first)
        sub = (slice(first - 1, None), slice(first - 1, None))
        for k in ("z1", "m1", "n1"):
            err = np.abs(snap[(L.id, k)][sub].astype(np.float64) - before[L.id][k][sub])
            assert err.max() <= 5.0e-6 * (1.0 + np.abs(before[L.id][k][sub]).max()), (L.id, k)
        assert not snap[(L.id, "zmin1")].any()
    assert np.abs(snap[(1, "zmax1")] - raw1).max() <= 5.0e-6 * (1.0 + np.abs(raw1).max())

    dev = DeviceRun(sc)
    init = {L.id: dev.state(L.id) for L in sc.layers}
    assert _read(2, start, tmp_path) == 0
    loaded = {}
    for L in sc.layers:
        first = 1 if L.id == 1 else 2
        g = dev.state(L.id)
        for k in ("z1", "m1", "n1"):
            exp = init[L.id][k].copy()
            exp[first - 1:, first - 1:] = snap[(L.id, k)][first - 1:, first - 1:]
            _same(g[k], exp, f"loaded layer {L.id} {k}")
            loaded[(L.id, k)] = exp
    _same(dev.maxima(1)["zmax_raw"], snap[(1, "zmax1")], "loaded raw maximum")
    dev.run(more)

    orc = Oracle(sc, 32, pow_mode=1)
    for L in sc.layers:
        st = {}
        for k in ("z", "m", "n"):
            st[k + "1"] = loaded[(L.id, k + "1")]
            st[k + "2"] = loaded[(L.id, k + "1")].copy()
        z = np.zeros_like(st["z1"])
        st["dz1"], st["dz2"] = z, z.copy()
        orc.set_state(L.id, st)
    orc.step(more)
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            _same(g[k], o[k], f"resumed layer {L.id} {k}")
        gm, om = dev.maxima(L.id), orc.maxima(L.id)
        _same(gm["zmax_wet"], om["zmax_wet"], f"resumed layer {L.id} wet maximum")
        exp_arr = np.where(om["arrival"] >= 0, om["arrival"] + done, -1)
        assert np.array_equal(gm["arrival"], exp_arr), f"resumed layer {L.id} arrival"
    dev.close()


def test_hotstart_type_1_and_missing_files(tmp_path):
    lib().gcomcot_set_error_mode_(ip(0))
    sc = _scenario()
    print("[gpu] hot start: START_TYPE 1 and a missing file", flush=True)
    dev = DeviceRun(sc)
    dev.run(9)
    _write(10, tmp_path)
    dev.close()
    dev = DeviceRun(sc)
    init2 = dev.state(2)
    assert _read(1, 10, tmp_path) == 0
    z1 = _file(tmp_path / "z1_01_000010.dat", (sc.layers[0].ny, sc.layers[0].nx), 1)
    _same(dev.state(1)["z1"], z1, "START_TYPE 1 layer 1")
    for k in ("z1", "m1", "n1"):
        _same(dev.state(2)[k], init2[k], f"START_TYPE 1 leaves layer 2 {k}")
    os.remove(tmp_path / "n1_02_000010.dat")
    assert _read(2, 10, tmp_path) == 1
    assert not dev.state(2)["n1"].any(), "a missing file zeroes its whole plane"
    # a malformed file is refused before any plane changes
    p = tmp_path / "m1_01_000010.dat"
    p.write_text("************" + p.read_text()[12:])
    before = dev.state(1)
    b = str(tmp_path).encode()
    miss = ctypes.c_int(0)
    assert lib().gcomcot_hotstart_read_(ip(2), ip(10), b, ip(len(b)), ctypes.byref(miss)) == 93
    assert last_error()[0] == 93  # the refusal is reported once
    for k in ("z1", "m1", "n1"):
        _same(dev.state(1)[k], before[k], f"refused load keeps layer 1 {k}")
    dev.close()
