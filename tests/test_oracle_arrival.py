"""Pins of the oracle's maximum and first-arrival accumulators.

The reference has no first-arrival field, and its only wet-test maximum is
commented out (output.f90:876-885); the GPU recorder it does run keeps a
maximum with no wet test, started from the initial elevation
(GPUcomcot.cu:79, GPUOutput.cu:46-53).  The oracle's max_arrival
(oracle.cpp) keeps all three after every (sub-)step of every layer.  Two
independent checks fix its definition:

1. Brute force.  The flang-built reference is driven through ALL_GRID's
   order one routine at a time (all_grids.f90:4-312), the elevation that
   each (sub-)step leaves is stored, and the three fields are derived by a
   scan over those snapshots: the raw maximum from Z0 with MAX_AMP's
   replacement test, the wet maximum from 0 over cells with H + Z2 > GX, and
   the root step (1-based count of completed root steps) of the first wet
   Z2 >= threshold.  Nested layers skip their ghost row I = 1 and column
   J = 1, which hold UPDATE's copies (all_grids.f90:323-324); the scan
   checks that those keep their initial values.  The oracle's own step
   must give the same planes as the driven reference and the same three
   fields as the scan, bit for bit.
2. Closed form.  On a flat Cartesian linear layer a raised plateau splits
   into two fronts of half its height moving at sqrt(g H).  With the
   threshold at a quarter of the plateau, the arrival step of a cell at
   distance r from the plateau's edge is r / (sqrt(g H) dt) plus a fixed
   start-up: MASS runs before MOMENT with M = 0, so the first step leaves
   Z unchanged (one step), an arrival is counted at the end of a whole
   step (under one), and the scheme's dispersion lets a sharp front lag by
   up to one more step over 100 cells.
"""
import math

import numpy as np
import pytest

from gcomcot import scenarios as S
from helpers import BEACH, bits
from oracle.bind import Oracle, Reference

GX = np.float32(1.0e-5)
OPS = dict(MASS=1, OPEN=2, MOMENT=3, JNQ=4, NEWQ=5, JNZ=6, UPDATE=7, CHANGE=9)


def _child(lc, lg, R, RT):
    return dict(laycord=lc, laygov=lg, footprint=(10, 14), rel_size=R, rel_time=RT, corner=(6, 8))


CASES = [
    ("root-nonlinear-shore", S.custom(laycord=0, laygov=1, nx=48, ny=40, **BEACH), 60),
    ("root-linear-cart", S.custom(laycord=1, laygov=0, nx=48, ny=40, cx=0.45), 60),
    ("nest-sph-sph-R2T2", S.custom(nx=48, ny=40, laygov=0, child=_child(0, 0, 2, 2), **BEACH), 30),
    ("nest-sph-cartNL-R3T4",
     S.custom(nx=48, ny=40, laygov=0, child=_child(1, 1, 3, 4), **BEACH), 30),
    ("nest-sph-sphNL-R3T5",
     S.custom(nx=48, ny=40, laygov=0, child=_child(0, 1, 3, 5), **BEACH), 24),
]


def _drive_reference(sc, steps):
    """Run the reference one routine at a time in ALL_GRID's order (one level
    of nesting) and return {layer id: [(root step, Z2 snapshot), ...]} with
    the elevation every (sub-)step leaves, plus the reference itself."""
    ref = Reference(sc)
    root = sc.layers[0]
    kids = [L for L in sc.layers if L.parent == root.id]
    snaps = {L.id: [] for L in sc.layers}
    for s in range(1, steps + 1):
        ref.call(OPS["MASS"], 1)
        ref.call(OPS["OPEN"], 1)
        for C in kids:
            nhalf = C.rel_time // 2 + 1
            for k in range(1, C.rel_time + 1):
                ref.call(OPS["JNQ"] if k == 1 else OPS["NEWQ"], C.id, k)
                ref.call(OPS["MASS"], C.id)
                ref.call(OPS["MOMENT"], C.id)
                snaps[C.id].append((s, ref.state(C.id)["z2"].copy()))
                if k == nhalf:
                    ref.call(OPS["JNZ"], C.id)
                ref.call(OPS["UPDATE"], C.id)
        ref.call(OPS["MOMENT"], 1)
        snaps[1].append((s, ref.state(1)["z2"].copy()))
        ref.call(OPS["CHANGE"], 1)
    return ref, snaps


def _scan(L, snaps, thr):
    """Raw maximum, wet maximum and first arrival by scanning the snapshots."""
    h = np.asarray(L.h, np.float32)
    raw = np.array(L.z0, np.float32, copy=True)
    wet_max = np.zeros_like(raw)
    arrival = np.full(raw.shape, -1, np.int32)
    live = np.ones(raw.shape, bool)
    if L.parent != -1:
        live[:, 0] = False   # I = 1
        live[0, :] = False   # J = 1
    for s, z in snaps:
        z = np.asarray(z, np.float32)
        upd = live & ~(raw >= z)
        raw = np.where(upd, z, raw)
        wet = live & ((h + z) > GX)
        first = wet & (z >= np.float32(thr)) & (arrival < 0)
        arrival = np.where(first, s, arrival)
        wet_max = np.where(wet & (z > wet_max), z, wet_max)
    return raw, wet_max, arrival, live


@pytest.mark.parametrize("thr", [0.01, 0.2], ids=["thr0.01", "thr0.2"])
@pytest.mark.parametrize("name,sc,steps", CASES, ids=[c[0] for c in CASES])
def test_accumulators_equal_a_scan_of_every_step(name, sc, steps, thr):
    if not Reference.available():
        pytest.skip("oracle/_ref (flang build of the reference) not built")
    ref, snaps = _drive_reference(sc, steps)
    # the driven sequence is the reference's own time loop
    whole = Reference(sc)
    whole.step(steps)
    orc = Oracle(sc, 32, pow_mode=0, arrival_threshold=thr)
    orc.step(steps)
    arrived = 0
    for L in sc.layers:
        a, b, o = ref.state(L.id), whole.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            assert np.array_equal(bits(a[k]), bits(b[k])), f"{name}: driven {k} of {L.id}"
            assert np.array_equal(bits(o[k]), bits(a[k])), f"{name}: oracle {k} of {L.id}"
        per = steps * (1 if L.parent == -1 else L.rel_time)
        assert len(snaps[L.id]) == per
        raw, wmax, arr, live = _scan(L, snaps[L.id], thr)
        got = orc.maxima(L.id)
        assert np.array_equal(bits(got["zmax_raw"]), bits(raw)), f"{name} layer {L.id} raw max"
        assert np.array_equal(bits(got["zmax_wet"]), bits(wmax)), f"{name} layer {L.id} wet max"
        assert np.array_equal(got["arrival"], arr), f"{name} layer {L.id} arrival"
        # ghosts keep the initial values
        assert np.all(got["arrival"][~live] == -1) and np.all(got["zmax_wet"][~live] == 0)
        arrived += int(np.count_nonzero(arr > 0))
        print(f"{name} thr {thr} layer {L.id}: {int(np.count_nonzero(arr > 0))} cells arrived, "
              f"steps {arr[arr > 0].min() if (arr > 0).any() else '-'}.."
              f"{arr.max()}", flush=True)
    assert arrived > 0, f"{name}: nothing arrived, the check is empty"


def test_arrival_front_moves_at_sqrt_gh():
    nx, ny, H, A, dxm = 360, 400, 100.0, 1.0, 500.0
    x0, x1 = 40, 150   # plateau on cells x0+1 .. x1 (1-based)
    h = np.full((ny, nx), H, np.float32)
    z0 = np.zeros((ny, nx), np.float32)
    z0[:, x0:x1] = A
    dt = S.cr_dt(1, 0, h, dxm, dxm, 0.0, 0.0)
    L = S.LayerSpec(id=1, parent=-1, level=1, laycord=1, laygov=0, nx=nx, ny=ny, dx=dxm,
                    dy=dxm, dt=dt, h=h, z0=z0, fric_switch=1, fric_coef=0.0)
    steps = 360
    orc = Oracle(S.Scenario("plateau", [L], steps), 32, pow_mode=0, arrival_threshold=A / 4)
    orc.step(steps)
    # the centre row lies outside the domain of dependence of the top and
    # bottom boundaries (200 rows at half a cell per step) for 400 steps
    arr = orc.maxima(1)["arrival"][ny // 2]
    c = math.sqrt(9.807 * H)
    right = [(i, ((i - 1) - (x1 - 0.5)) * dxm) for i in range(x1 + 4, 330, 12)]
    left = [(i, ((x0 - 0.5) - (i - 1)) * dxm) for i in range(x0 - 4, 8, -8)]
    offs = {}
    for side, pts in (("right", right), ("left", left)):
        for i, r in pts:
            t = r / (c * dt)
            assert arr[i - 1] > 0, f"{side} cell {i} never reached A/4"
            off = arr[i - 1] - t
            offs[(side, round(r))] = off
            assert 1.0 - 1e-3 <= off <= 3.0 + 1e-3, \
                f"{side} cell {i}: arrival {arr[i-1]}, r/(c dt) = {t:.2f}"
    # both fronts see the same start-up
    for (side, r), off in offs.items():
        if side == "left" and ("right", r) in offs:
            assert abs(off - offs[("right", r)]) <= 1.0
    print("arrival - r/(c dt):", {k: round(v, 2) for k, v in offs.items()})
