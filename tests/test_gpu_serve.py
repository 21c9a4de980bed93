"""GPU: the continuous-batching request loop.

Requests arrive on a Poisson clock faster than they complete, under a
footprint budget that holds only a few of them, so they queue, are
admitted as others finish and run interleaved in the batch.  Every
forecast (wet maximum and first-arrival fields) must equal, bit for bit,
the oracle's run of that scenario alone.
"""
import numpy as np
import pytest

from gcomcot.serve import Request, Server, poisson_arrivals, summarize
from helpers import diff_report, small_scenarios
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

CASES = small_scenarios()


def test_served_forecasts_equal_runs_alone():
    steps = [60 + 7 * i for i in range(len(CASES))]
    t = poisson_arrivals(400.0, len(CASES), seed=5)
    reqs = [Request(i, t[i], sc, steps[i]) for i, (_, sc, _) in enumerate(CASES)]
    budget = sum(sorted(r.footprint for r in reqs)[-3:])
    print(f"[gpu] serving {len(reqs)} requests, budget {budget} B", flush=True)
    srv = Server(budget_bytes=budget, chunk=16, keep_results=True)
    out = srv.serve(reqs)
    assert srv.peak_bytes <= budget
    assert srv.peak_running >= 2  # requests overlapped on the device
    for r in out:
        name = CASES[r.rid][0]
        assert r.arrival <= r.admitted <= r.done, name
        orc = Oracle(r.scenario, 32, pow_mode=1)
        orc.step(r.steps)
        om = orc.maxima(1)
        nb, mx = diff_report(r.result["zmax_wet"], om["zmax_wet"])
        assert nb == 0, f"{name}: served max differs from alone in {nb} cells (max {mx})"
        assert np.array_equal(r.result["arrival"], om["arrival"]), f"{name}: arrival differs"
    s = summarize(out)
    print(f"[gpu] p50 {s['p50_ttf_s']:.4f} s p99 {s['p99_ttf_s']:.4f} s", flush=True)
    srv.close()


def test_request_over_budget_is_refused():
    _, sc, _ = CASES[0]
    r = Request(0, 0.0, sc, 5)
    srv = Server(budget_bytes=r.footprint - 1)
    with pytest.raises(ValueError):
        srv.serve([r])
    srv.close()
