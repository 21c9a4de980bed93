"""GPU: fault sources computed on the device.

The library evaluates DEFORM_OKADA in FP64 on the device; the oracle
(oracle/okada.py, pinned against the flang-built reference in
tests/test_okada_oracle.py) evaluates the same FP64 formula on the CPU.
Both round the final sum to REAL once, so the fields agree bit for bit
except where the last FP64 bits of the two math libraries straddle a REAL
rounding boundary (one float ulp) or the GX cut.  The spreading onto
nested layers (CELL_INTERP) and the updates Z + D, H - D are REAL
arithmetic and must match the oracle bit for bit on the device's own
root field; the step afterwards must see the new depth exactly as a
scenario built with it.
"""
import copy

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot.batch import Batch
from helpers import BEACH, diff_report
from oracle import okada as OK
from oracle import outfiles as OF
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


def _faults(xc, yc):
    return [OK.Fault(10000, 100000, 50000, 5, 30, 15, 90, x0=xc, y0=yc),
            OK.Fault(5000, 60000, 30000, 3, 200, 45, 30, x0=xc, y0=yc, switch=9),
            OK.Fault(20000, 150000, 60000, 8, 10, 80, 0, x0=xc, y0=yc)]


def _assert_fp64_twins(dev, orc, tag):
    dev = np.asarray(dev, np.float32)
    orc = np.asarray(orc, np.float32)
    ulp = np.spacing(np.abs(orc))
    d = np.abs(dev.astype(np.float64) - orc.astype(np.float64))
    far = d > np.maximum(ulp, 2 * OK.GX).astype(np.float64)
    assert not far.any(), f"{tag}: {far.sum()} cells beyond one ulp / the GX cut, max {d.max():.3e}"
    same = np.count_nonzero(dev.view(np.int32) == orc.view(np.int32))
    assert same >= 0.999 * dev.size, f"{tag}: only {same} of {dev.size} cells bitwise equal"


@pytest.mark.parametrize("kind", ["sph-2min", "sph-0.5min", "cart"])
def test_device_okada_matches_the_oracle(kind):
    if kind == "cart":
        sc = S.custom(laycord=1, laygov=0, nx=160, ny=128)
    else:
        sc = S.custom(laycord=0, laygov=0, nx=160, ny=128, dx=2.0 if kind == "sph-2min" else 0.5)
    L = sc.layers[0]
    x, y = S.layer_coords(L)
    faults = _faults(140.4, 10.3) if kind == "cart" else _faults(float(x[80]), float(y[64]))
    print(f"[gpu] okada {kind}: {len(faults)} faults on {L.nx}x{L.ny}", flush=True)
    for f in faults:
        if kind == "cart":
            f.xo, f.yo = 140.0, 10.0
        dev = DeviceRun(sc)
        dev.set_coords()
        dev.fault(f.params(), f.switch)
        got = dev.deform(1)
        dev.close()
        want = OK.deform_okada(f, L.nx, L.ny, L.laycord, x, y, L.dx, L.dy)
        assert np.abs(want).max() > 0.01
        _assert_fp64_twins(got, want, f"{kind} fault {f.th}/{f.dl}/{f.rd}")


def _nest():
    child = dict(laycord=0, laygov=1, footprint=(20, 24), rel_size=3, rel_time=3, corner=(30, 26))
    return S.custom(laycord=0, laygov=1, nx=96, ny=80, dx=2.0, child=child, **BEACH)


def test_deformation_spreads_and_updates_every_layer_exactly():
    sc = _nest()
    x, y = S.layer_coords(sc.layers[0])
    f = _faults(float(x[45]), float(y[38]))[0]
    print("[gpu] fault on a nest", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    dev.fault(f.params(), f.switch)
    d_root = dev.deform(1)
    for L in sc.layers:
        d = dev.deform(L.id)
        if L.id != 1:
            want = OK.cell_interp(d_root, (L.ny, L.nx), L.corners, L.rel_size)
            nb, mx = diff_report(d, want)
            assert nb == 0, f"layer {L.id} CELL_INTERP: {nb} cells differ (max {mx})"
        z0 = np.asarray(L.z0, np.float32)
        h0 = np.asarray(L.h, np.float32)
        np.testing.assert_array_equal(dev.state(L.id)["z1"], z0 + d)
        np.testing.assert_array_equal(dev.depth(L.id), h0 - d)
    np.testing.assert_array_equal(dev.maxima(1)["zmax_raw"], dev.state(1)["z1"])
    dev.close()


def test_steps_after_the_source_use_the_new_depth():
    sc = _nest()
    x, y = S.layer_coords(sc.layers[0])
    f = _faults(float(x[45]), float(y[38]))[1]
    n = 40
    print(f"[gpu] fault then {n} steps", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    dev.fault(f.params(), f.switch)
    moved = copy.deepcopy(sc)
    for L in moved.layers:
        d = dev.deform(L.id)
        L.h = np.asarray(L.h, np.float32) - d
        L.z0 = np.asarray(L.z0, np.float32) + d
    dev.run(n)
    orc = Oracle(moved, 32, pow_mode=1)
    orc.step(n)
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
    dev.close()


def test_batched_scenario_with_a_source_equals_the_single_run():
    sc = _nest()
    x, y = S.layer_coords(sc.layers[0])
    f = _faults(float(x[45]), float(y[38]))[2]
    n = 30
    print(f"[gpu] batched fault, {n} steps", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    dev.fault(f.params(), f.switch)
    dev.run(n)
    alone = {L.id: dev.state(L.id) for L in sc.layers}
    dev.close()
    b = Batch()
    other = b.add(S.custom(laycord=1, laygov=1, nx=64, ny=56, **BEACH), n)
    sid = b.add(sc, n)
    b.fault(sid, f.params(), f.switch)
    b.run(n)
    for L in sc.layers:
        got = b.state(sid, L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(got[k], alone[L.id][k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
    assert b.progress(other) == (n, n)
    b.close()


def _firing_call(dt, t0, kmax):
    """The GET_FLOOR_DEFORM call at which a rupture at t0 applies: 0 before
    the loop (TIME = 0), k before root step k (TIME after step k, REAL);
    deform.f90:29, comcot.f90:323, 341."""
    times = OF.model_times(dt, kmax)
    t0 = np.float32(t0)
    for k, t in enumerate(times):
        if t0 >= t and t0 < np.float32(t + np.float32(dt)):
            return k
    return None


def test_scheduled_ruptures_apply_at_the_reference_steps():
    sc = _nest()
    x, y = S.layer_coords(sc.layers[0])
    dt = np.float32(sc.layers[0].dt)
    faults = _faults(float(x[45]), float(y[38]))
    t0s = [np.float32(0.5) * dt, np.float32(7.5) * dt, np.float32(12.0) * dt]
    n = 20
    calls = [_firing_call(dt, t, n) for t in t0s]
    assert calls[0] == 0 and all(c is not None for c in calls)
    print(f"[gpu] scheduled ruptures at calls {calls}", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    for f, t in zip(faults, t0s):
        dev.schedule_fault(f.params(), f.switch, float(t))
    dev.run(8)
    dev.run(n - 8, graph=False)
    assert dev.pending_faults() == 0
    sched = {L.id: dev.state(L.id) for L in sc.layers}
    dev.close()
    dev = DeviceRun(sc)   # the same by hand: each fault before its step
    dev.set_coords()
    done = 0
    for f, k in sorted(zip(faults, calls), key=lambda p: p[1]):
        if k - 1 > done:
            dev.run(k - 1 - done)
            done = k - 1
        dev.fault(f.params(), f.switch)
    dev.run(n - done)
    for L in sc.layers:
        st = dev.state(L.id)
        for key in ("z1", "m1", "n1"):
            nb, mx = diff_report(sched[L.id][key], st[key])
            assert nb == 0, f"layer {L.id} {key}: {nb} cells differ (max {mx})"
    dev.close()


def test_fault_errors_are_reported():
    from gcomcot._lib import GcomcotError
    sc = S.custom(laycord=0, laygov=0, nx=48, ny=40)
    dev = DeviceRun(sc)
    f = _faults(1.0, 11.0)[0]
    with pytest.raises(GcomcotError, match="coordinates"):
        dev.fault(f.params(), 0)          # spherical root without coordinates
    dev.set_coords()
    with pytest.raises(GcomcotError, match="switch"):
        dev.fault(f.params(), 3)
    bad = f.params()
    bad[3] = np.nan
    with pytest.raises(GcomcotError, match="finite"):
        dev.fault(bad, 0)
    dev.close()
