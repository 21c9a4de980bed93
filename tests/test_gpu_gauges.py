"""GPU: gauge time series sampled on the device.

The reference's own gauge routine GET_TS (output.f90:1035-1155), flang-built
into oracle/_ref with a CPU GCOMCOT_GET_Z, samples the reference's state
after every root step; the library samples its own state on the device in
the same step loop.  On linear layers (no pow/exp) the two states are
bitwise equal, so every sample must be too: stencil search, weights and
the interpolation order are all pinned by that comparison, including
gauges outside the grid (clamped stencils) and gauges on a nested layer.
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from oracle.bind import Reference

pytestmark = pytest.mark.gpu


def _points(sc):
    pts = []
    for L in sc.layers:
        x, y = S.layer_coords(L)
        rng = np.random.default_rng(L.id)
        for _ in range(5):
            pts.append((float(rng.uniform(x[1], x[-2])), float(rng.uniform(y[1], y[-2])), L.id))
        pts.append((float(x[0]) - 1.0, float(y[-1]) + 1.0, L.id))      # outside: clamped
        pts.append((float(x[3]), float(y[4]), L.id))                   # on a node
    return pts


@pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")
@pytest.mark.parametrize("case", ["sph-linear", "cart-linear", "nest-linear"])
def test_gauges_match_the_reference_get_ts(case):
    child = dict(laycord=0, laygov=0, footprint=(14, 12), rel_size=3, rel_time=3, corner=(10, 9))
    sc = {"sph-linear": S.custom(laycord=0, laygov=0, nx=56, ny=48, cx=0.45),
          "cart-linear": S.custom(laycord=1, laygov=0, nx=56, ny=48, cx=0.45),
          "nest-linear": S.custom(laycord=0, laygov=0, nx=56, ny=48, cx=0.45, child=child)}[case]
    steps = 60
    pts = _points(sc)
    print(f"[gpu] gauges {case}: {len(pts)} gauges, {steps} steps", flush=True)
    ref = Reference(sc)
    for L in sc.layers:
        ref.set_coords(L.id, *S.layer_coords(L))
    want = np.zeros((len(pts), steps), np.float32)
    for k in range(steps):
        ref.step(1)
        for g, (x0, y0, lid) in enumerate(pts):
            want[g, k] = ref.get_ts(x0, y0, lid)
    dev = DeviceRun(sc)
    dev.set_gauges(pts, capacity=steps)
    dev.run(steps // 2, graph=True)
    dev.run(steps - steps // 2, graph=False)
    got = dev.gauges(steps)
    dev.close()
    assert got.shape == want.shape
    bad = np.argwhere(got.view(np.int32) != want.view(np.int32))
    assert bad.size == 0, f"{len(bad)} samples differ, first {bad[:3]}: {got[tuple(bad[0])]} vs {want[tuple(bad[0])]}"


def test_gauge_ring_keeps_the_last_samples():
    sc = S.custom(laycord=0, laygov=1, nx=48, ny=40, cx=0.45)
    x, y = S.layer_coords(sc.layers[0])
    pts = [(float(x[20]) + 0.003, float(y[15]) + 0.002, 1)]
    dev = DeviceRun(sc)
    dev.set_gauges(pts, capacity=16)
    dev.run(40)
    ring = dev.gauges(64)
    assert ring.shape == (1, 16)
    # the newest sample is the interpolated current surface
    st = dev.state(1)["z1"]
    cx = np.float32((np.float32(pts[0][0]) - x[20]) / (x[21] - x[20]))
    cy = np.float32((np.float32(pts[0][1]) - y[15]) / (y[16] - y[15]))
    one = np.float32(1.0)
    z = (st[15, 20] * (one - cx) * (one - cy) + st[15, 21] * cx * (one - cy)
         + st[16, 20] * (one - cx) * cy + st[16, 21] * cx * cy)
    assert ring[0, -1] == np.float32(z)
    dev.close()


@pytest.mark.parametrize("case", ["sph", "cart", "nest-child"])
def test_gauges_reproduce_an_affine_surface(case):
    """GET_TS (output.f90:1035-1155) interpolates bilinearly in the layer's
    own coordinates, which reproduces an affine surface z = a + b x + c y
    exactly up to rounding, wherever the gauge lies in its stencil cell.
    After the first root step the interior elevation is still the initial
    one (continuity reads the level-1 fluxes, which start at 0), so every
    sample must equal the affine function at the gauge position: a shifted
    stencil, swapped or mirrored weights, or x and y transposed all miss
    by a sizeable fraction of the surface's change over one cell."""
    child = dict(laycord=1, laygov=0, footprint=(14, 12), rel_size=3, rel_time=1, corner=(10, 9))
    sc = {"sph": S.custom(laycord=0, laygov=0, nx=56, ny=48),
          "cart": S.custom(laycord=1, laygov=0, nx=56, ny=48),
          "nest-child": S.custom(laycord=0, laygov=0, nx=56, ny=48, child=child)}[case]
    rng = np.random.default_rng(3)
    lids = [2] if case == "nest-child" else [1]
    coef = {}
    for L in sc.layers:
        L.h = np.full((L.ny, L.nx), 3000.0, np.float32)
        x, y = S.layer_coords(L)
        X, Y = np.meshgrid(x.astype(np.float64), y.astype(np.float64))
        # about 0.3 m of change across a cell in each direction
        b, c = 0.3 / float(x[1] - x[0]), -0.2 / float(y[1] - y[0])
        a = -(b * float(x[L.nx // 2]) + c * float(y[L.ny // 2]))
        coef[L.id] = (a, b, c)
        L.z0 = (a + b * X + c * Y).astype(np.float32)
    pts = []
    for lid in lids:
        L = sc.layers[lid - 1]
        x, y = S.layer_coords(L)
        lo = 2 if lid == 1 else 3  # stencils off the boundary ring and ghosts
        for _ in range(40):
            pts.append((float(rng.uniform(x[lo], x[-3])), float(rng.uniform(y[lo], y[-3])), lid))
        pts.append((float(x[7]), float(y[5]), lid))  # on a node
    print(f"[gpu] gauges on an affine surface, {case}: {len(pts)} gauges", flush=True)
    dev = DeviceRun(sc)
    dev.set_gauges(pts, capacity=4)
    dev.run(1, graph=False)
    got = dev.gauges(4)[:, 0].astype(np.float64)
    z_int = {lid: dev.state(lid)["z1"] for lid in lids}
    dev.close()
    for lid in lids:
        L = sc.layers[lid - 1]
        lo = 1 if lid == 1 else 2
        assert np.array_equal(z_int[lid][lo:-1, lo:-1], L.z0[lo:-1, lo:-1]), \
            "the first step moved the interior surface"
    for g, (x0, y0, lid) in enumerate(pts):
        a, b, c = coef[lid]
        want = a + b * float(np.float32(x0)) + c * float(np.float32(y0))
        scale = np.abs(sc.layers[lid - 1].z0).max()
        assert abs(got[g] - want) <= 2e-6 * scale, \
            f"gauge {g} at ({x0}, {y0}) on layer {lid}: {got[g]} vs {want}"
