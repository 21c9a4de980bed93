"""Early-warning request loop over the batch engine (continuous batching).

Requests (a Scenario plus its number of root steps) arrive on a clock.
The loop admits waiting requests in arrival order while the summed device
footprint of the admitted ones stays under a budget, advances every
admitted scenario together in chunks of root steps (scenarios join and
leave the running batch between chunks; their state stays resident on the
device), and completes a request once its maximum-elevation and arrival
fields are on the host.  Time-to-forecast = completion - arrival.
"""
import collections
import time

import numpy as np

from .batch import Batch, scenario_footprint


class Request:
    __slots__ = ("rid", "arrival", "scenario", "steps", "admitted", "done", "sid", "result",
                 "footprint")

    def __init__(self, rid, arrival, scenario, steps=None):
        self.rid = rid
        self.arrival = float(arrival)
        self.scenario = scenario
        self.steps = scenario.steps if steps is None else int(steps)
        self.admitted = None
        self.done = None
        self.sid = None
        self.result = None
        self.footprint = scenario_footprint(scenario)


def poisson_arrivals(rate, n, seed):
    """n arrival times (s) of a Poisson process of `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class Server:
    """Continuous-batching loop; one per process (it drives the library's
    single batch engine)."""

    def __init__(self, budget_bytes=250 * 10 ** 9, chunk=128, keep_results=False,
                 arrival_threshold=0.01):
        self.budget = int(budget_bytes)
        self.chunk = int(chunk)
        self.keep = keep_results
        self.batch = Batch(arrival_threshold=arrival_threshold)
        self.peak_bytes = 0     # largest admitted footprint sum seen
        self.peak_running = 0   # most scenarios on the device at once

    def serve(self, requests, clock=time.perf_counter):
        """Run until every request has completed; returns the requests with
        their admitted/done times (seconds from the start)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.footprint > self.budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, over the budget")
        pending = collections.deque(reqs)
        waiting = collections.deque()
        running = {}
        used = 0
        t0 = clock()
        while pending or waiting or running:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting and used + waiting[0].footprint <= self.budget:
                r = waiting.popleft()
                r.sid = self.batch.add(r.scenario, r.steps)
                r.admitted = clock() - t0
                running[r.sid] = r
                used += r.footprint
                self.peak_bytes = max(self.peak_bytes, used)
                self.peak_running = max(self.peak_running, len(running))
            if not running:
                if pending:  # idle until the next request arrives
                    time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            self.batch.run(self.chunk)
            for sid in self.batch.finished():
                r = running.pop(sid)
                mx = self.batch.maxima(sid, 1)  # the forecast: max-eta and arrival fields
                r.done = clock() - t0
                if self.keep:
                    r.result = mx
                self.batch.remove(sid)
                used -= r.footprint
        return reqs

    def close(self):
        self.batch.close()


def summarize(reqs):
    ttf = np.array([r.done - r.arrival for r in reqs])
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return {"requests": len(reqs), "p50_ttf_s": float(np.percentile(ttf, 50)),
            "p99_ttf_s": float(np.percentile(ttf, 99)), "max_ttf_s": float(ttf.max()),
            "scenarios_per_s": len(reqs) / span if span > 0 else float("inf"),
            "span_s": span}
