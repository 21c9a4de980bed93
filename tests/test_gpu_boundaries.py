"""GPU: sponge layer and wall boundary in the device loop, bit for bit.

The sponge acts inside the momentum pass of layer 1 (SPONGE_LAYER,
boundaries.f90:206-246, which the reference's GPU build does not run);
the oracle (pinned against the flang-built reference's own SPONGE_COEF and
SPONGE_LAYER in tests/test_boundaries_oracle.py) must give the same bits
after every step, with the coefficients the library computes from the
layer's depth and coordinates.  A wall (BC_WALL) is a depth transform made
before the layers are created; the loop must treat the walled border
exactly as the oracle does.
"""
import copy

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from helpers import BEACH, diff_report
from oracle import sponge as SP
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "cart-nonlinear": dict(laycord=1, laygov=1, nx=96, ny=80, **BEACH),
    "cart-linear": dict(laycord=1, laygov=0, nx=96, ny=80, cx=0.45),
    "cart-nest": dict(laycord=1, laygov=0, nx=96, ny=80, **BEACH,
                      child=dict(laycord=1, laygov=1, footprint=(20, 24), rel_size=3,
                                 rel_time=3, corner=(30, 26))),
}


def _compare(dev, orc, sc, tag):
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"{tag} layer {L.id} {k}: {nb} cells differ (max {mx})"
        gm, om = dev.maxima(L.id), orc.maxima(L.id)
        assert diff_report(gm["zmax_wet"], om["zmax_wet"])[0] == 0
        assert np.array_equal(gm["arrival"], om["arrival"])


@pytest.mark.parametrize("case", sorted(CASES))
def test_sponge_matches_the_oracle(case):
    sc = S.custom(**CASES[case])
    R = sc.layers[0]
    x, y = S.layer_coords(R)
    n = 70
    print(f"[gpu] sponge {case}, {n} steps", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords()
    dev.sponge(True)
    dev.run(n // 2)
    dev.run(n - n // 2, graph=False)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.set_sponge(*SP.sponge_coef(R.laycord, R.h, x, y))
    orc.step(n)
    _compare(dev, orc, sc, case)
    dev.close()


def test_walled_layer_matches_the_oracle():
    sc = S.custom(laycord=1, laygov=1, nx=96, ny=80, **BEACH)
    walled = copy.deepcopy(sc)
    walled.layers[0].h = S.bc_wall(sc.layers[0].h)
    n = 60
    print(f"[gpu] wall, {n} steps", flush=True)
    dev = DeviceRun(walled)
    dev.run(n)
    orc = Oracle(walled, 32, pow_mode=1)
    orc.step(n)
    _compare(dev, orc, walled, "wall")
    g = dev.state(1)
    assert not g["m1"][:, :2].any() and not g["n1"][:2, :].any()
    dev.close()
