"""The hot-start text files against the flang-built reference.

The reference writes and reads its hot-start snapshots with HOTSTART_DATA
(hotstart.f90:30-166): per layer five files z1_LL_SSSSSS.dat, m1_, n1_,
zmax1_, zmin1_, one '(10F12.5)' WRITE per row J over I, J >= 1 (outermost
layer) or >= 2 (nested layers).  The library's host formatter and reader
(gcomcot_hotstart_file_write_ / _read_, the ones its device entry points
use) must give the reference's bytes when writing and the reference's
REAL values when reading:

* writing: a two-layer nest after a few steps, with planted values whose
  F12.5 form is delicate (exact ties at the fifth decimal, -0.0 and small
  negatives that print as -0.00000, values too wide for twelve characters,
  NaN and infinities), byte for byte against HOT_START with START_TYPE 0;
* reading: files in forms the Fortran F editing accepts beyond what it
  writes (implied decimals, exponents with E or D, blank and short
  records), loaded by HOT_START with START_TYPE 2 and by the library,
  bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

from gcomcot import scenarios as S
from gcomcot._lib import ip, lib
from helpers import BEACH, bits
from oracle.bind import Reference

pytestmark = pytest.mark.skipif(not Reference.available(), reason="reference not built")

KINDS = ("z1", "m1", "n1", "zmax1", "zmin1")
STEP = 37


def _nest():
    return S.custom(nx=48, ny=40, laygov=0, child=dict(laycord=1, laygov=1, footprint=(10, 14),
                                                       rel_size=3, rel_time=4, corner=(6, 8)),
                    **BEACH)


def _lib_write(path, plane, first):
    ny, nx = plane.shape
    a = np.ascontiguousarray(plane, np.float32)
    b = path.encode()
    rc = lib().gcomcot_hotstart_file_write_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                            ip(nx), ip(ny), ip(first))
    assert rc == 0


def _lib_read(path, shape, first, fill=np.float32(7.0)):
    a = np.full(shape, fill, np.float32)
    b = path.encode()
    rc = lib().gcomcot_hotstart_file_read_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                           ip(shape[1]), ip(shape[0]), ip(first))
    assert rc == 0, rc
    return a


def _planes(ref, lid):
    st = ref.state(lid)
    zmax, zmin = ref.zmaxmin(lid)
    return dict(z1=st["z1"], m1=st["m1"], n1=st["n1"], zmax1=zmax, zmin1=zmin)


def test_writer_gives_the_reference_bytes(tmp_path, monkeypatch):
    sc = _nest()
    ref = Reference(sc)
    ref.step(12)
    rng = np.random.default_rng(5)
    odd = np.array([0.015625, -0.015625, 0.046875, -0.0, -1e-7, 1e-7, 123456.7, 999999.99,
                    1.0e6, -99999.9, -999999.9, 5e-6, 2.5e-6, 7.5e-6, 1.2345675, np.nan,
                    np.inf, -np.inf, 3.0e38, -2.4e-39], np.float32)
    for L in sc.layers:
        st = ref.state(L.id)
        z = st["z1"].copy()
        flat = z.reshape(-1)
        flat[5:5 + odd.size] = odd
        flat[-odd.size:] = odd[::-1]
        st["z1"] = z
        st["m1"] = (st["m1"] + rng.normal(0, 50.0, z.shape)).astype(np.float32)
        ref.set_state(L.id, st)
        zmax = rng.normal(0.0, 3.0, z.shape).astype(np.float32)
        zmin = -np.abs(rng.normal(0.0, 300.0, z.shape)).astype(np.float32)
        ref.zmaxmin(L.id, zmax, zmin)
    a, b = tmp_path / "ref", tmp_path / "lib"
    a.mkdir()
    b.mkdir()
    monkeypatch.chdir(a)
    ref.hot_start(0, STEP)
    want = sorted(f"{k}_{L.id:02d}_{STEP:06d}.dat" for L in sc.layers for k in KINDS)
    assert sorted(os.listdir(a)) == want
    for L in sc.layers:
        first = 1 if L.id == 1 else 2
        for k, plane in _planes(ref, L.id).items():
            name = f"{k}_{L.id:02d}_{STEP:06d}.dat"
            _lib_write(str(b / name), plane, first)
            x, y = (a / name).read_bytes(), (b / name).read_bytes()
            if x != y:
                k0 = next((i for i, (u, v) in enumerate(zip(x, y)) if u != v), min(len(x), len(y)))
                raise AssertionError(f"{name}: differs from byte {k0}: "
                                     f"{x[k0 - 20:k0 + 20]!r} vs {y[k0 - 20:k0 + 20]!r}")
    assert b"************" in (a / f"z1_01_{STEP:06d}.dat").read_bytes()
    assert b"         NaN" in (a / f"z1_01_{STEP:06d}.dat").read_bytes()


def _field(v, form):
    if form == "plain":
        s = f"{v:.5f}"
    elif form == "implied":  # no decimal point: five implied decimals
        s = str(int(round(v * 1e5)))
    elif form == "E":
        s = f"{v:.4E}"
    elif form == "D":
        s = f"{v:.3E}".replace("E", "D")
    elif form == "short":
        s = f"{v:.2f}"
    else:
        s = "" if v == 0 else f"{v:.5f}"
    return s.rjust(12)[-12:]


def _hand_file(path, plane, first, rng):
    ny, nx = plane.shape
    lines = []
    for j in range(first - 1, ny):
        row = plane[j, first - 1:]
        for i0 in range(0, row.size, 10):
            chunk = row[i0:i0 + 10]
            fields = [_field(float(v), rng.choice(["plain", "implied", "E", "D", "short",
                                                   "blank"])) for v in chunk]
            text = "".join(fields)
            if rng.random() < 0.1:
                text = text.rstrip()  # a short record reads as blank-padded
            lines.append(text)
    path.write_text("\n".join(lines) + "\n")


def test_reader_gives_the_reference_values(tmp_path, monkeypatch):
    sc = _nest()
    ref = Reference(sc)
    rng = np.random.default_rng(11)
    planes = {}
    for L in sc.layers:
        first = 1 if L.id == 1 else 2
        for k in KINDS:
            p = rng.normal(0.0, 2.0, (L.ny, L.nx)).astype(np.float32)
            p[rng.random(p.shape) < 0.05] = 0.0
            name = f"{k}_{L.id:02d}_{STEP:06d}.dat"
            _hand_file(tmp_path / name, p, first, rng)
            planes[(L.id, k)] = p
    monkeypatch.chdir(tmp_path)
    ref.hot_start(2, STEP)
    for L in sc.layers:
        first = 1 if L.id == 1 else 2
        got = _planes(ref, L.id)
        for k in KINDS:
            mine = _lib_read(str(tmp_path / f"{k}_{L.id:02d}_{STEP:06d}.dat"), (L.ny, L.nx),
                             first)
            sub = (slice(first - 1, None), slice(first - 1, None))
            assert np.array_equal(bits(mine[sub]), bits(got[k][sub])), f"layer {L.id} {k}"
            if first == 2:  # the first row and column are not in the file
                assert (mine[0, :] == 7.0).all() and (mine[:, 0] == 7.0).all()


def test_reader_refuses_asterisks_and_short_files(tmp_path):
    lib().gcomcot_set_error_mode_(ip(0))     # errors return codes instead of exiting
    p = np.zeros((4, 12), np.float32)
    good = tmp_path / "ok.dat"
    _lib_write(str(good), p, 1)
    text = good.read_text().splitlines()
    bad = tmp_path / "stars.dat"
    bad.write_text("\n".join([("*" * 12) + text[0][12:]] + text[1:]) + "\n")
    b = str(bad).encode()
    a = np.zeros_like(p)
    assert lib().gcomcot_hotstart_file_read_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                             ip(12), ip(4), ip(1)) == 93
    short = tmp_path / "short.dat"
    short.write_text("\n".join(text[:-1]) + "\n")
    b = str(short).encode()
    assert lib().gcomcot_hotstart_file_read_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                             ip(12), ip(4), ip(1)) == 92
    b = str(tmp_path / "none.dat").encode()
    assert lib().gcomcot_hotstart_file_read_(b, ip(len(b)), a.ctypes.data_as(ctypes.c_void_p),
                                             ip(12), ip(4), ip(1)) == 96
