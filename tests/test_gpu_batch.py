"""GPU: the scenario batch engine, bit for bit.

Every scenario of a batch must end exactly where the oracle (device-math
pow/exp) ends after that scenario's own number of root steps, and where
the same scenario ends when run alone: the batch only changes which
launch a tile rides in, never its arithmetic.  Scenario step counts differ
(odd and even) so plans are rebuilt as scenarios finish, and one test adds
a scenario to a batch that is already running (continuous batching).
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot.batch import Batch, scenario_footprint
from helpers import diff_report, small_scenarios
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

CASES = small_scenarios()


def _oracle(sc, steps):
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(steps)
    return orc


def _same(a, b, tag):
    nb, mx = diff_report(a, b)
    assert nb == 0, f"{tag}: {nb} cells differ (max {mx})"


def _check_vs_oracle(b, sid, sc, orc, tag):
    for L in sc.layers:
        g, o = b.state(sid, L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            _same(g[k], o[k], f"{tag} layer {L.id} {k}")
        gm, om = b.maxima(sid, L.id), orc.maxima(L.id)
        _same(gm["zmax_wet"], om["zmax_wet"], f"{tag} layer {L.id} zmax")
        assert np.array_equal(gm["arrival"], om["arrival"]), f"{tag} layer {L.id} arrival"


@pytest.mark.parametrize("two_step,graphs,streams,fused",
                         [(True, True, 3, True), (False, True, 3, True), (True, False, 3, True),
                          (True, True, 2, True), (True, True, 1, True), (True, True, 3, False),
                          (True, False, 1, False)],
                         ids=["two-step", "one-step", "two-step-no-graphs", "two-streams",
                              "one-stream", "separate-halves", "separate-halves-no-graphs"])
def test_mixed_batch_matches_oracle(two_step, graphs, streams, fused):
    """Single-layer and nested scenarios together; with more than one stream
    the single-layer ones run beside the nested chain(s); nested linear roots
    with their momentum launch computing the next continuity or not."""
    steps = [s + (i % 3) for i, (_, _, s) in enumerate(CASES)]  # odd and even counts
    print(f"[gpu] batch of {len(CASES)} scenarios, steps {steps}, two_step={two_step}, "
          f"graphs={graphs}, streams={streams}, fused_halves={fused}", flush=True)
    b = Batch(two_step=two_step, graphs=graphs, streams=streams, fused_halves=fused)
    sids = [b.add(sc, n) for (_, sc, _), n in zip(CASES, steps)]
    assert b.run(10 ** 6) == len(CASES)
    for (name, sc, _), sid, n in zip(CASES, sids, steps):
        print(f"[gpu]   check {name} ({n} steps)", flush=True)
        assert b.progress(sid) == (n, n)
        _check_vs_oracle(b, sid, sc, _oracle(sc, n), name)
    info = b.info()
    assert info["plans"] >= 2 and (info["graph_replays"] > 0) == graphs, info
    b.close()


def test_streamed_stores_in_the_batch():
    """The store policy changed after the first plan was built: the plan is
    rebuilt (the batch kernels store through the caches whatever the
    policy), and every scenario still lands on the oracle's bits."""
    from gcomcot._lib import lib, ip
    steps = [s + (i % 3) for i, (_, _, s) in enumerate(CASES)]
    print(f"[gpu] batch with streamed stores, steps {steps}", flush=True)
    b = Batch()
    sids = [b.add(sc, n) for (_, sc, _), n in zip(CASES, steps)]
    b.run(7)
    plans = b.info()["plans"]
    assert lib().gcomcot_set_store_policy_(ip(2)) == 0
    try:
        assert b.run(10 ** 6) == len(CASES)
        assert b.info()["plans"] > plans
    finally:
        lib().gcomcot_set_store_policy_(ip(1))
    for (name, sc, _), sid, n in zip(CASES, sids, steps):
        _check_vs_oracle(b, sid, sc, _oracle(sc, n), f"{name} streamed")
    b.close()


@pytest.mark.parametrize("shape", range(7))
def test_two_step_shapes_in_the_batch(shape):
    """Every two-step tile shape on single-layer linear roots, with step
    counts that leave 1, 2 and 3 steps after the last four-step plan, beside
    a nest and a nonlinear layer that take one step per launch."""
    picks = [CASES[0], CASES[1], CASES[2], CASES[5]]
    steps = [41, 42, 43, 44]
    print(f"[gpu] batch two-step shape {shape}", flush=True)
    b = Batch(two_step_shape=shape)
    sids = [b.add(sc, n) for (_, sc, _), n in zip(picks, steps)]
    b.run(10 ** 6)
    for (name, sc, _), sid, n in zip(picks, sids, steps):
        _check_vs_oracle(b, sid, sc, _oracle(sc, n), f"{name} two-step shape {shape}")
    b.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_every_batch_tile_shape(shape):
    picks = [CASES[2], CASES[3], CASES[6], CASES[7]]
    print(f"[gpu] batch tile shape {shape}", flush=True)
    b = Batch(autotune=False, shape=shape)
    sids = [b.add(sc, 41) for _, sc, _ in picks]
    b.run(41)
    for (name, sc, _), sid in zip(picks, sids):
        _check_vs_oracle(b, sid, sc, _oracle(sc, 41), f"{name} shape {shape}")
    b.close()


def test_alone_and_in_mix_identical():
    """A scenario gives the same bits alone (drop-in loop, batch of one) and
    admitted into a running batch of others."""
    name, sc, _ = CASES[7]
    n = 57
    print(f"[gpu] alone vs mix: {name}", flush=True)
    dev = DeviceRun(sc)
    dev.run(n)
    alone = {L.id: dev.state(L.id) for L in sc.layers}
    alone_max = {L.id: dev.maxima(L.id) for L in sc.layers}
    dev.close()

    b = Batch()
    sid = b.add(sc, n)
    b.run(n)
    one = {L.id: b.state(sid, L.id) for L in sc.layers}
    b.close()

    b = Batch()
    others = [b.add(o, 90 + i) for i, (_, o, _) in enumerate(CASES) if o is not sc]
    b.run(23)  # the mix is under way when the request arrives
    sid = b.add(sc, n)
    b.run(10 ** 6)
    for L in sc.layers:
        mix = b.state(sid, L.id)
        mm = b.maxima(sid, L.id)
        for k in ("z1", "m1", "n1"):
            _same(mix[k], alone[L.id][k], f"mix vs alone layer {L.id} {k}")
            _same(one[L.id][k], alone[L.id][k], f"batch-of-one vs alone layer {L.id} {k}")
        _same(mm["zmax_wet"], alone_max[L.id]["zmax_wet"], f"mix vs alone layer {L.id} zmax")
        assert np.array_equal(mm["arrival"], alone_max[L.id]["arrival"])
    for o in others:
        assert b.progress(o)[0] == b.progress(o)[1]
    b.close()


def test_footprint_accounting_and_removal():
    b = Batch()
    sids = [b.add(sc, 5) for _, sc, _ in CASES[:4]]
    want = sum(scenario_footprint(sc) for _, sc, _ in CASES[:4])
    assert b.device_bytes() == want
    b.run(5)
    assert sorted(b.finished()) == sorted(sids)
    for s in sids:
        b.remove(s)
    assert b.device_bytes() == 0 and b.finished() == []
    b.close()
