"""GPU: the benchmark configurations at the sizes the oracle can follow.

* The nonlinear single-layer generator (friction, moving shorelines,
  dipole source) at 512 x 384 for 2000 steps against the reference's own
  FP32 arithmetic: the oracle with libm pow/exp, which the golden fixtures
  pin to the flang build of the Fortran bit for bit.  The device uses its
  own deterministic pow/exp, so here the test is the stated tolerance, not
  bit equality: wet/dry masks and first-arrival steps identical, elevation
  within 1e-5 m and fluxes within 1e-4 m^2/s after the full run.  The error
  of both against the FP64 oracle is printed.
* The linear single layer at its full size, 1024 x 768 for 2000 steps, bit
  for bit.
* The two-level nest with its refinement pairing (REL_SIZE 4, REL_TIME 4,
  Cartesian nonlinear child over the shelf ramp) at reduced size, bit for
  bit.
* The scenario mix through the batch engine at reduced sizes, with
  children of REL_SIZE 3, 4 and 5 and REL_TIME from CR_CHECK, each
  scenario bit for bit.
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot.batch import Batch
from helpers import compare_device_oracle, diff_report
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

GX = np.float32(1.0e-5)
TOL_Z, TOL_MN = 1.0e-5, 1.0e-4


def _oracles(sc, steps, runs):
    """Oracle runs [(precision, pow_mode), ...] of `steps` root steps."""
    out = []
    for prec, pm in runs:
        o = Oracle(sc, prec, pow_mode=pm)
        o.step(steps)
        out.append(o)
    return out


def _wet(z, h):
    return (np.asarray(z, np.float32) + np.asarray(h, np.float32)) >= GX


def test_nonlinear_512x384_2000_steps_within_tolerance_of_reference_fp32():
    sc = S.nonlinear_layer(nx=512, ny=384, dx=2.0)
    steps = 2000
    print(f"[gpu] {sc.name}: {steps} steps, device vs reference FP32 arithmetic", flush=True)
    dev = DeviceRun(sc)
    dev.run(steps)
    g, gm = dev.state(1), dev.maxima(1)
    dev.close()
    ref, f64 = _oracles(sc, steps, [(32, 0), (64, 0)])
    o, om, d = ref.state(1), ref.maxima(1), f64.state(1)
    h = sc.layers[0].h
    wg, wo = _wet(g["z1"], h), _wet(o["z1"], h)
    assert np.array_equal(wg, wo), f"wet/dry masks differ in {np.count_nonzero(wg != wo)} cells"
    assert np.array_equal(gm["arrival"], om["arrival"]), \
        f"arrival steps differ in {np.count_nonzero(gm['arrival'] != om['arrival'])} cells"
    arrived = int(np.count_nonzero(om["arrival"] >= 0))
    assert arrived > 10000, f"only {arrived} cells reached the arrival threshold"
    err = {}
    for k, tol in (("z1", TOL_Z), ("m1", TOL_MN), ("n1", TOL_MN)):
        a, b, c = (np.asarray(x[k], np.float64) for x in (g, o, d))
        err[k] = (float(np.abs(a - b).max()), float(np.abs(a - c).max()),
                  float(np.abs(b - c).max()))
        assert err[k][0] <= tol, f"{k}: max |device - reference| = {err[k][0]} > {tol}"
    nb, mx = diff_report(gm["zmax_wet"], om["zmax_wet"])
    assert mx <= TOL_Z, f"wet maximum differs by {mx}"
    for k, (dr, df, rf) in err.items():
        print(f"[gpu]   {k}: |device-ref32| {dr:.3g}  |device-fp64| {df:.3g}  "
              f"|ref32-fp64| {rf:.3g}", flush=True)
    print(f"[gpu]   wet cells {int(wg.sum())}, arrived {arrived}, zmax cells differing {nb}",
          flush=True)


def test_linear_1024x768_full_size_bit_exact():
    sc = S.linear_layer()
    assert (sc.layers[0].nx, sc.layers[0].ny, sc.steps) == (1024, 768, 2000)
    print(f"[gpu] {sc.name}: {sc.steps} steps, bit for bit", flush=True)
    dev = DeviceRun(sc)
    dev.run(sc.steps)
    (orc,) = _oracles(sc, sc.steps, [(32, 1)])
    compare_device_oracle(sc, dev, orc, sc.name)
    arrived = int(np.count_nonzero(orc.maxima(1)["arrival"] >= 0))
    assert arrived > 100000, f"only {arrived} cells reached the arrival threshold"
    dev.close()


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_two_level_nest_r4_t4_reduced(graph):
    sc = S.two_level_nest(pnx=256, pny=192, footprint=40, rel_size=4, rel_time=4, corner_i=20)
    C = sc.layers[1]
    assert (C.rel_size, C.rel_time, C.laycord, C.laygov) == (4, 4, 1, 1)
    steps = 240
    print(f"[gpu] {sc.name} R4T4 graph={graph}: {steps} steps", flush=True)
    dev = DeviceRun(sc)
    dev.run(steps, graph=graph)
    (orc,) = _oracles(sc, steps, [(32, 1)])
    compare_device_oracle(sc, dev, orc, sc.name)
    assert np.count_nonzero(orc.maxima(2)["arrival"] >= 0) > 0, "the wave never reaches the child"
    dev.close()


def test_scenario_mix_reduced_through_the_batch_engine():
    mix = S.scenario_mix(n=8, case=1, min_cells=96 * 72, max_cells=256 * 192, nest_share=0.5,
                         steps=(60, 140), child_side=(60, 140))
    pairs = sorted({(L.rel_size, L.rel_time) for sc in mix for L in sc.layers[1:]})
    assert {r for r, _ in pairs} == {3, 4, 5}, pairs
    print(f"[gpu] scenario mix of {len(mix)} through the batch engine, children {pairs}",
          flush=True)
    b = Batch()
    sids = [b.add(sc) for sc in mix]
    assert b.run(10 ** 6) == len(mix)
    orcs = [_oracles(sc, sc.steps, [(32, 1)])[0] for sc in mix]
    for sc, sid, orc in zip(mix, sids, orcs):
        print(f"[gpu]   check {sc.name} ({sc.steps} steps)", flush=True)
        for L in sc.layers:
            gs, os_ = b.state(sid, L.id), orc.state(L.id)
            for k in ("z1", "m1", "n1"):
                nb, mx = diff_report(gs[k], os_[k])
                assert nb == 0, f"{sc.name} layer {L.id} {k}: {nb} differ (max {mx})"
            gm, om = b.maxima(sid, L.id), orc.maxima(L.id)
            nb, mx = diff_report(gm["zmax_wet"], om["zmax_wet"])
            assert nb == 0, f"{sc.name} layer {L.id} zmax: {nb} differ"
            assert np.array_equal(gm["arrival"], om["arrival"]), f"{sc.name} layer {L.id} arrival"
    b.close()
