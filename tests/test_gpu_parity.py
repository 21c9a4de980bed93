"""GPU parity: the HIP step against the oracle, bit for bit.

The oracle runs in device-math mode (pow/exp of detmath.h, the only place
the kernels may differ from libm); everything else is the restated
reference arithmetic, so planes, wet/dry masks, maxima and arrival steps
must match exactly.  Small grids keep the oracle side to seconds.
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from helpers import compare_device_oracle, diff_report, small_scenarios
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

CASES = small_scenarios()


def _compare(sc, dev, orc, tag):
    compare_device_oracle(sc, dev, orc, tag)


@pytest.mark.parametrize("name,sc,steps", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_device_loop_matches_oracle(name, sc, steps, graph):
    print(f"[gpu] {name} graph={graph} steps={steps}", flush=True)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(steps)
    dev = DeviceRun(sc)
    dev.run(steps, graph=graph)
    _compare(sc, dev, orc, name)
    dev.close()


@pytest.mark.parametrize("name,sc,steps", CASES, ids=[c[0] for c in CASES])
def test_every_tile_shape(name, sc, steps):
    """Every tile variant of the fused step (tiles and strips), single layers
    and the nested layers of every nest pairing (the autotuner may give a
    nested layer any of them)."""
    from gcomcot._lib import lib, ip
    import ctypes
    n = ctypes.c_int(0)
    lib().gcomcot_tile_variants_(ctypes.byref(n), None, None)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(40)
    try:
        for v in range(n.value):
            print(f"[gpu] {name} tile variant {v}", flush=True)
            assert lib().gcomcot_set_tile_variant_(ip(v)) == 0
            dev = DeviceRun(sc)
            dev.run(40, graph=False)
            _compare(sc, dev, orc, f"{name} variant {v}")
            dev.close()
    finally:
        lib().gcomcot_set_tile_variant_(ip(0))


NESTS = [c for c in CASES if c[0].startswith("nest")]


@pytest.mark.parametrize("name,sc,steps", NESTS, ids=[c[0] for c in NESTS])
@pytest.mark.parametrize("fused", [True, False], ids=["fused-halves", "separate-halves"])
def test_nest_root_halves_fused_or_separate(name, sc, steps, fused):
    """A linear root with nested layers, its momentum launch also computing
    the next root step's continuity or not, over calls of 1, 2, 7 and 8
    steps with and without graphs, timed steps and a drop-in step between
    them: every call ends on a whole step, so the mix lands on the oracle."""
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(37)
    dev = DeviceRun(sc)
    dev.set_fused_halves(fused)
    try:
        for n, graph in [(1, True), (2, False), (7, True), (8, True), (3, False)]:
            print(f"[gpu] {name} fused={fused} run {n} graph={graph}", flush=True)
            dev.run(n, graph=graph)
        print(f"[gpu] {name} fused={fused} timed 5 + drop-in 1 + run 10", flush=True)
        dev.timed(5)
        dev.dropin_step()
        dev.run(10, graph=True)
        _compare(sc, dev, orc, f"{name} fused={fused}")
    finally:
        dev.set_fused_halves(True)
        dev.close()


@pytest.mark.parametrize("name,sc,steps", [CASES[0], CASES[2], CASES[3], CASES[5], CASES[6]],
                         ids=[c[0] for c in (CASES[0], CASES[2], CASES[3], CASES[5], CASES[6])])
@pytest.mark.parametrize("policy", [2, 0], ids=["streamed", "cached"])
def test_time_level_store_policies(name, sc, steps, policy):
    """Every layer's new time level written with nontemporal stores (what
    the smallest and the largest layers do by default) or through the
    caches (what layers in between do), through the device loop with
    graphs (two steps per launch on the linear single layer), the
    momentum + next continuity launches of the nests and the drop-in entry
    points: the oracle's bits."""
    from gcomcot._lib import GcomcotError
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(steps)
    dev = DeviceRun(sc)
    with pytest.raises(GcomcotError, match="code 46"):
        dev.set_store_policy(3)
    dev.set_store_policy(policy)
    try:
        print(f"[gpu] {name} store policy {policy}", flush=True)
        dev.run(steps - 1, graph=True)
        dev.dropin_step()
        _compare(sc, dev, orc, f"{name} store policy {policy}")
    finally:
        dev.set_store_policy(1)
        dev.close()


def test_reference_entry_points_match_oracle():
    """The Fortran driver's call sequence through the reference symbols,
    with layers registered by gcomcot_init_layer_ from full PARA planes."""
    sc = S.custom(laycord=0, laygov=0, nx=72, ny=64, cx=0.45,
                  child=dict(laycord=0, laygov=0, footprint=(20, 36), rel_size=2, rel_time=2,
                             corner=(8, 14)))
    steps = 60
    print("[gpu] drop-in reference entry points", flush=True)
    orc = Oracle(sc, 32, pow_mode=0)
    orc.step(steps)
    dev = DeviceRun(sc, reference_init=True)
    for _ in range(steps):
        dev.dropin_step()
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"layer {L.id} {k}: {nb} differ (max {mx})"
    nb, mx = diff_report(dev.maxima(1)["zmax_raw"], orc.maxima(1)["zmax_raw"])
    assert nb == 0, f"raw max: {nb} differ"
    dev.close()


def test_linear_layer_reduced_matches_oracle():
    sc = S.linear_layer(nx=256, ny=192)
    print("[gpu] linear layer 256x192 x 400 steps", flush=True)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(400)
    dev = DeviceRun(sc)
    dev.run(400)
    _compare(sc, dev, orc, "linear-256")
    dev.close()


def test_nonlinear_layer_reduced_matches_oracle():
    sc = S.nonlinear_layer(nx=256, ny=192, dx=2.0)
    print("[gpu] nonlinear layer 256x192 x 300 steps", flush=True)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(300)
    dev = DeviceRun(sc)
    dev.run(300)
    _compare(sc, dev, orc, "nonlinear-256")
    dev.close()


LINEAR = [CASES[0], CASES[1],
          ("ragged-linear-77x61", S.custom(laycord=0, laygov=0, nx=77, ny=61, cx=0.45), 90)]


@pytest.mark.parametrize("mode", range(8), ids=["single", "2step-64x8", "2step-32x16",
                                                 "2step-64x16", "2step-64x16-512",
                                                 "2step-64x16-1024", "2step-32x16-512",
                                                 "2step-64x32-1024"])
def test_linear_layers_one_or_two_steps_per_launch(mode):
    """Linear single layers run two root steps per launch by default; every
    tile shape of that kernel and the single-step kernel must give the
    oracle's bits, for odd and even step counts, with and without graphs."""
    from gcomcot._lib import ip, lib
    lib().gcomcot_set_temporal_blocking_(ip(mode))
    try:
        for name, sc, steps in LINEAR:
            for n, graph in ((steps + 1, True), (steps, False)):
                print(f"[gpu] {name} blocking={mode} steps={n} graph={graph}", flush=True)
                orc = Oracle(sc, 32, pow_mode=1)
                orc.step(n)
                dev = DeviceRun(sc)
                dev.run(n, graph=graph)
                _compare(sc, dev, orc, f"{name} mode {mode}")
                dev.close()
    finally:
        lib().gcomcot_set_temporal_blocking_(ip(1))


@pytest.mark.parametrize("name,sc,steps", [CASES[0], CASES[2], CASES[4], CASES[6], CASES[7]],
                         ids=[CASES[0][0], CASES[2][0], CASES[4][0], CASES[6][0], CASES[7][0]])
def test_autotune_mid_run_keeps_the_state(name, sc, steps):
    """The library autotuner times dry launches on the live layer: a run
    tuned half way must end exactly where the oracle does."""
    from gcomcot._lib import ip, lib
    print(f"[gpu] autotune mid-run: {name}", flush=True)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(77)
    dev = DeviceRun(sc)
    try:
        dev.run(37)
        v, b, ms = dev.autotune(reps=2)
        assert ms > 0 and v >= 0
        assert 0 <= b <= 7
        if sc.layers[0].laygov != 0 and len(sc.layers) == 1:
            assert b == 0  # two steps per launch are for linear layers only
        dev.run(40)
        _compare(sc, dev, orc, f"{name} tuned")
    finally:
        dev.close()
        lib().gcomcot_set_tile_variant_(ip(0))
        lib().gcomcot_set_temporal_blocking_(ip(1))
