"""The oracle of the output files, pinned by the reference's own writers.

The flang-built reference (oracle/_ref) runs its DATA_PRT and MAX_AMP
(output.f90:758-1031) on its own stepped states and writes real files into
a temporary directory; oracle/outfiles.py must give the same names and the
same bytes from the same states.  That pins the record layout, the file
names (including Fortran's asterisks for a step count that does not fit),
the dry-land mask and the tide correction, and the half-hour / end-of-run
schedule of the maximum files over a sweep of model times.
"""
import os

import numpy as np
import pytest

from gcomcot import scenarios as S
from helpers import BEACH
from oracle import outfiles as OF
from oracle.bind import Reference

needs_ref = pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")


def _files(d):
    out = {}
    for n in os.listdir(d):
        with open(os.path.join(d, n), "rb") as f:
            out[n] = f.read()
    return out


def _nest():
    child = dict(laycord=0, laygov=1, footprint=(20, 36), rel_size=3, rel_time=3, corner=(8, 14))
    return S.custom(nx=72, ny=64, laygov=1, child=child, **BEACH)


@needs_ref
@pytest.mark.parametrize("tide", [0.0, 0.75])
def test_snapshot_files_match_data_prt(tmp_path, monkeypatch, tide):
    sc = _nest()
    flux = {1: 0, 2: 1}
    ref = Reference(sc)
    monkeypatch.chdir(tmp_path)
    want = {}
    dry_seen = wet_land_seen = False
    for k in range(0, 61):
        if k:
            ref.step(1)
        if k % 20:
            continue
        for L in sc.layers:
            ref.data_prt(k, L.id, flux[L.id], tide)
            st = ref.state(L.id)
            lvl = "1" if k == 0 else "2"  # K = 0 writes the host level 1, later CUDA_GETZ's level 2
            z, m, n = st["z" + lvl], st["m" + lvl], st["n" + lvl]
            want.update(OF.snapshot_files(L.id, k, flux[L.id], z, m, n, L.h, tide))
            land = L.h < 0
            dry = land & (z + L.h <= OF.GX)
            dry_seen |= bool(dry.any())
            wet_land_seen |= bool((land & ~dry).any())
    assert dry_seen and wet_land_seen, "the case must have dry and flooded land"
    got = _files(tmp_path)
    assert sorted(got) == sorted(want)
    for name, body in want.items():
        assert got[name] == body, name
    assert "m_01_000020.dat" in got and "m_02_000020.dat" not in got


@needs_ref
def test_step_count_beyond_six_digits_gives_asterisks(tmp_path, monkeypatch):
    sc = S.custom(laycord=1, laygov=0, nx=24, ny=20)
    ref = Reference(sc)
    monkeypatch.chdir(tmp_path)
    ref.data_prt(1234567, 1, 1, 0.0)
    assert sorted(os.listdir(tmp_path)) == ["z_01_******.dat"]
    assert OF.snapshot_names(1, 1234567, 1) == ["z_01_******.dat"]
    assert OF.snapshot_names(3, 999999, 0) == ["z_03_999999.dat", "m_03_999999.dat",
                                               "n_03_999999.dat"]
    assert OF.snapshot_names(2, 5, 2) == []


@needs_ref
@pytest.mark.parametrize("dx,tide", [(2.0, 0.0), (0.75, -0.4)])
def test_maximum_files_match_max_amp(tmp_path, monkeypatch, dx, tide):
    sc = S.custom(laycord=0, laygov=1, nx=40, ny=32, dx=dx, **BEACH)
    L = sc.layers[0]
    ref = Reference(sc)
    for _ in range(25):  # a maximum with structure: wet, dry and flooded cells
        ref.step(1)
        ref.maxamp()
    zraw = ref.raw_max()
    monkeypatch.chdir(tmp_path)
    dt = np.float32(L.dt)
    kend = int(np.ceil(2.2 * 3600.0 / float(dt)))
    tend = np.float32(kend * dt)
    times = OF.model_times(dt, kend + 2)
    written = []
    for k in range(1, kend + 3):
        ref.max_amp(times[k], tend, tide)
        names = os.listdir(tmp_path)
        want = OF.maxamp_name(times[k], dt, tend)
        assert names == ([want] if want else []), (k, float(times[k]), names, want)
        if want:
            with open(os.path.join(tmp_path, want), "rb") as f:
                assert f.read() == OF.maxamp_file(zraw, L.h, tide), (k, want)
            os.remove(os.path.join(tmp_path, want))
            written.append(want)
    halves = [w for w in written if w != "zmax_layer01.dat"]
    assert len(halves) >= 4 and written[-1] == "zmax_layer01.dat"
    sched = OF.schedule(dt, tend, 0, 2, 1, kend + 2)
    assert [z for _, _, z in sched] == written


def test_shown_elevation_cases():
    """The mask on hand-picked cells: sea, dry land, flooded land, the GX
    edge, and the tide correction that only applies above GX."""
    z = np.array([0.5, 0.0, 3.0, -2.0 + 1e-6, 0.25], np.float32)
    h = np.array([10.0, -1.0, -2.0, 2.0, -0.25], np.float32)
    # cell 4: Z + H = 0 <= GX and H < 0 -> dry; cell 3: H > 0 -> never masked
    np.testing.assert_array_equal(OF.shown_elevation(z, h), [0.5, 0.0, 3.0, z[3], 0.0])
    np.testing.assert_array_equal(OF.shown_elevation(z, h, 1.0),
                                  np.array([1.5, 0.0, 4.0, z[3] + 1, 0.0], np.float32))
    np.testing.assert_array_equal(OF.shown_elevation(z, h, 5e-6), OF.shown_elevation(z, h))


def test_fortran_record_layout():
    a = np.arange(6, dtype=np.float32).reshape(2, 3)  # (ny=2, nx=3)
    rec = OF.fortran_record(a)
    assert rec[:4] == rec[-4:] == (24).to_bytes(4, "little")
    np.testing.assert_array_equal(np.frombuffer(rec[4:-4], "<f4"), np.arange(6))
