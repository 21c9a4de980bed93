"""GPU: exact checkpoints of the device state.

A run saved after an odd number of root steps (so the time-level parity
is flipped) and restored into a fresh layer tree must continue to the
same bits as the run that was never interrupted: planes of both time
levels, the wet and reference maxima and the arrival field.  A checkpoint
is refused by a tree with different layers before anything changes.  A
batched scenario saved, removed and resumed later in a new slot finishes
exactly where an uninterrupted batch finishes it.
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot._lib import GcomcotError
from gcomcot.batch import Batch
from helpers import BEACH, diff_report

pytestmark = pytest.mark.gpu


def _nest():
    child = dict(laycord=0, laygov=1, footprint=(20, 24), rel_size=3, rel_time=3, corner=(30, 26))
    return S.custom(laycord=0, laygov=1, nx=96, ny=80, dx=2.0, child=child, **BEACH)


def _snapshot(dev, sc):
    out = {}
    for L in sc.layers:
        st = dev.state(L.id)
        mx = dev.maxima(L.id)
        out[L.id] = dict(st, **mx)
    return out


def _same(a, b, tag):
    for lid in a:
        for k in a[lid]:
            if k == "arrival":
                assert np.array_equal(a[lid][k], b[lid][k]), f"{tag} layer {lid} arrival"
            else:
                nb, mx = diff_report(a[lid][k], b[lid][k])
                assert nb == 0, f"{tag} layer {lid} {k}: {nb} cells differ (max {mx})"


def test_restored_run_continues_to_the_same_bits(tmp_path):
    sc = _nest()
    ck = tmp_path / "run.ckpt"
    print("[gpu] checkpoint after 37 steps, 23 more", flush=True)
    dev = DeviceRun(sc)
    dev.run(37)
    dev.save(ck)
    dev.run(23)
    want = _snapshot(dev, sc)
    dev.close()
    dev = DeviceRun(sc)
    dev.run(5)                    # a different state, overwritten by the load
    dev.load(ck)
    dev.run(23)
    _same(want, _snapshot(dev, sc), "restored")
    dev.close()


def test_checkpoint_of_other_layers_is_refused(tmp_path):
    sc = _nest()
    ck = tmp_path / "run.ckpt"
    dev = DeviceRun(sc)
    dev.run(3)
    dev.save(ck)
    dev.close()
    other = S.custom(laycord=0, laygov=1, nx=96, ny=80, dx=2.0, **BEACH)
    dev = DeviceRun(other)
    dev.run(2)
    before = dev.state(1)["z1"]
    with pytest.raises(GcomcotError, match="layers"):
        dev.load(ck)
    np.testing.assert_array_equal(dev.state(1)["z1"], before)
    (tmp_path / "junk").write_bytes(b"not a checkpoint")
    with pytest.raises(GcomcotError, match="not a checkpoint"):
        dev.load(tmp_path / "junk")
    dev.close()


def test_truncated_checkpoint_is_refused_before_anything_changes(tmp_path):
    sc = _nest()
    ck = tmp_path / "run.ckpt"
    print("[gpu] truncated checkpoint", flush=True)
    dev = DeviceRun(sc)
    dev.run(5)
    dev.save(ck)
    body = ck.read_bytes()
    ck.write_bytes(body[:len(body) - 1000])   # the last layer record cut short
    dev.run(4)
    before = _snapshot(dev, sc)
    with pytest.raises(GcomcotError, match="truncated"):
        dev.load(ck)
    _same(_snapshot(dev, sc), before, "after a refused load")
    dev.run(3)                                # and the loop goes on
    dev.close()


def test_batched_scenario_preempted_and_resumed(tmp_path):
    sc = _nest()
    side = S.custom(laycord=1, laygov=1, nx=64, ny=56, **BEACH)
    n = 40
    print("[gpu] batch preemption", flush=True)
    b = Batch()
    ref_sid = b.add(sc, n)
    b.add(side, n)
    b.run(n)
    want = {L.id: b.state(ref_sid, L.id) for L in sc.layers}
    b.close()

    b = Batch()
    sid = b.add(sc, n)
    other = b.add(side, n)
    b.run(17)
    ck = tmp_path / "req.ckpt"
    b.save(sid, ck)
    b.remove(sid)
    b.run(9)                      # the rest of the batch goes on meanwhile
    again = b.add(sc, n)
    b.load(again, ck)
    assert b.progress(again) == (17, n)
    b.run(10 ** 6)
    for L in sc.layers:
        got = b.state(again, L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(got[k], want[L.id][k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
    assert b.progress(other) == (n, n)
    b.close()


def test_restore_keeps_only_the_ruptures_still_to_come(tmp_path):
    """Scheduled faults are not part of a checkpoint: after a load, those whose
    GET_FLOOR_DEFORM call lies at or before the restored step are dropped
    (their deformation is in the state), later ones still fire."""
    from oracle import okada as OK
    sc = _nest()
    ck = tmp_path / "run.ckpt"
    x, y = S.layer_coords(sc.layers[0])
    dt = np.float32(sc.layers[0].dt)
    early = OK.Fault(10000, 100000, 50000, 5, 30, 15, 90, x0=float(x[45]), y0=float(y[38]))
    late = OK.Fault(5000, 60000, 30000, 3, 200, 45, 30, x0=float(x[45]), y0=float(y[38]),
                    switch=9)
    times = (float(np.float32(2.5) * dt), float(np.float32(17.5) * dt))
    print("[gpu] checkpoint with scheduled ruptures", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    for f, t in zip((early, late), times):
        dev.schedule_fault(f.params(), f.switch, t)
    dev.run(9)
    dev.save(ck)
    dev.run(16)
    assert dev.pending_faults() == 0
    want = _snapshot(dev, sc)
    dev.close()
    dev = DeviceRun(sc)
    dev.set_coords()
    for f, t in zip((early, late), times):
        dev.schedule_fault(f.params(), f.switch, t)
    assert dev.pending_faults() == 2
    dev.load(ck)
    assert dev.pending_faults() == 1
    dev.run(16)
    _same(_snapshot(dev, sc), want, "restored with a rupture to come")
    dev.close()
