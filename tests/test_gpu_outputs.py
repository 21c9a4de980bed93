"""GPU: reference-format output files written by the device loop.

The library packs the files on the device, copies them out on a stream of
its own and writes them from a thread while the step loop goes on.  Every
file must have the name and the bytes that the oracle's restatement of
DATA_PRT / MAX_AMP (oracle/outfiles.py, itself pinned byte for byte by the
flang-built reference's writers) gives for the oracle's state at the same
root step, and the set of files must be the same.  Runs are cut into calls
of odd lengths, with and without graphs, so that output steps fall inside
graph replays, on two-step launches and across call boundaries.
"""
import os

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from helpers import BEACH, diff_report
from oracle import outfiles as OF
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "nest":
        child = dict(laycord=0, laygov=1, footprint=(20, 36), rel_size=3, rel_time=5, corner=(8, 14))
        sc = S.custom(nx=72, ny=64, laygov=0, child=child, **BEACH)
        return sc, dict(steps=70, iprt=9, out_opt=2, tend_steps=68, tide=0.5,
                        flux={1: 0, 2: 0}, chunks=[(23, True), (20, False), (27, True)])
    if name == "sc-nest":
        from helpers import sc_nest
        from oracle.bind import Reference
        if not Reference.available():
            pytest.skip("the SC nest is laid out by the flang-built reference")
        sc, _ = sc_nest()
        return sc, dict(steps=40, iprt=7, out_opt=2, tend_steps=38, tide=0.0,
                        flux={1: 0, 2: 0}, chunks=[(13, True), (27, False)])
    if name == "nonlinear":
        sc = S.custom(laycord=0, laygov=1, nx=72, ny=64, **BEACH)
        return sc, dict(steps=70, iprt=10, out_opt=2, tend_steps=200, tide=0.0,
                        flux={1: 1}, chunks=[(35, True), (35, True)])
    # linear single layer: two root steps per launch in the device loop
    sc = S.custom(laycord=0, laygov=0, nx=72, ny=64, cx=0.45)
    return sc, dict(steps=60, iprt=5, out_opt=0, tend_steps=58, tide=0.0,
                    flux={1: 0}, chunks=[(17, True), (43, True)])


def _expected(sc, c):
    dt = np.float32(sc.layers[0].dt)
    tend = np.float32(c["tend_steps"] * dt)
    due = {k: (snap, z) for k, snap, z in OF.schedule(dt, tend, c["iprt"], c["out_opt"], 0,
                                                       c["steps"])}
    orc = Oracle(sc, 32, pow_mode=1)
    files, writes = {}, 0
    for k in range(c["steps"] + 1):
        if k:
            orc.step(1)
        if k not in due:
            continue
        snap, zname = due[k]
        if snap:
            for L in sc.layers:
                st = orc.state(L.id)
                f = OF.snapshot_files(L.id, k, c["flux"].get(L.id, 1), st["z1"], st["m1"],
                                      st["n1"], L.h, c["tide"])
                files.update(f)
                writes += len(f)
        if zname:
            files[zname] = OF.maxamp_file(orc.maxima(1)["zmax_raw"], sc.layers[0].h, c["tide"])
            writes += 1
    return files, writes, tend, orc


@pytest.mark.parametrize("name", ["nest", "sc-nest", "nonlinear", "linear-two-step"])
def test_output_files_match_the_oracle(tmp_path, name):
    sc, c = _case(name)
    print(f"[gpu] output files: {name}, {c['steps']} steps", flush=True)
    want, writes, tend, orc = _expected(sc, c)
    dev = DeviceRun(sc)
    dev.set_output(tmp_path, c["iprt"], c["out_opt"], tend, c["tide"], c["flux"])
    for n, graph in c["chunks"]:
        dev.run(n, graph=graph)
    nfiles = dev.flush_output()
    raw = dev.maxima(1)["zmax_raw"]
    dev.close()
    got = {}
    for f in os.listdir(tmp_path):
        with open(os.path.join(tmp_path, f), "rb") as fh:
            got[f] = fh.read()
    assert sorted(got) == sorted(want)
    for f, body in want.items():
        if got[f] != body:
            a = np.frombuffer(got[f][4:-4], "<f4")
            b = np.frombuffer(body[4:-4], "<f4")
            nb, mx = diff_report(a, b) if a.size == b.size else (-1, -1)
            pytest.fail(f"{f}: bytes differ ({nb} cells, max {mx})")
    assert nfiles == writes
    if c["out_opt"] in (0, 2):
        nb, mx = diff_report(raw, orc.maxima(1)["zmax_raw"])
        assert nb == 0, f"recorder maximum: {nb} cells differ (max {mx})"


def test_unwritable_directory_is_reported_at_flush(tmp_path):
    """A write that fails does not stop the loop; the flush reports it."""
    from gcomcot._lib import GcomcotError
    sc, c = _case("nonlinear")
    print("[gpu] outputs into a missing directory", flush=True)
    dev = DeviceRun(sc)
    dev.set_output(tmp_path / "no" / "such" / "dir", 10, 1, 1000.0, 0.0, {1: 1})
    dev.run(25)
    z = dev.state(1)["z1"]
    with pytest.raises(GcomcotError, match="cannot open"):
        dev.flush_output()
    dev.close()
    ref = DeviceRun(sc)                       # the step results are those of a run without files
    ref.run(25)
    np.testing.assert_array_equal(ref.state(1)["z1"].view(np.int32), z.view(np.int32))
    ref.close()


def test_outputs_off_leave_the_loop_as_it_was(tmp_path):
    """Without gcomcot_output_set_ nothing is written and the recorder
    maximum stays at the initial elevation (only maxamp_launch_ raises it)."""
    sc, c = _case("nonlinear")
    print("[gpu] outputs off", flush=True)
    dev = DeviceRun(sc)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        dev.run(12)
        assert dev.flush_output() == 0
    finally:
        os.chdir(cwd)
    raw = dev.maxima(1)["zmax_raw"]
    dev.close()
    assert os.listdir(tmp_path) == []
    assert np.array_equal(raw.view(np.int32), np.asarray(sc.layers[0].z0, np.float32).view(np.int32))
