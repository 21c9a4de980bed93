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
        """keep_results: True keeps every forecast on its Request (result),
        a set of request ids keeps those only."""
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
                if self.keep is True or (self.keep and r.rid in self.keep):
                    r.result = mx
                self.batch.remove(sid)
                used -= r.footprint
        return reqs

    def close(self):
        self.batch.close()


def summarize(reqs, warmup=0.1):
    """Time-to-forecast statistics of a served stream.  Over all requests:
    p50/p99/max and scenarios/s from the first arrival to the last
    completion.  Steady state: the requests after the first `warmup` share
    of arrivals (the queue filling up) and before the arrivals stop; the
    offered rate is sustained when completions keep pace with arrivals
    (throughput over that window >= 90% of the offered rate) and the
    time-to-forecast does not climb (median of the window's last quarter
    below twice that of its first quarter)."""
    rs = sorted(reqs, key=lambda r: r.arrival)
    ttf = np.array([r.done - r.arrival for r in rs])
    arr = np.array([r.arrival for r in rs])
    done = np.array([r.done for r in rs])
    span = done.max() - arr.min()
    n = len(rs)
    k0 = int(round(warmup * n))
    win = slice(k0, n)
    offered = (n - 1) / (arr[-1] - arr[0]) if n > 1 and arr[-1] > arr[0] else float("inf")
    t_lo, t_hi = arr[k0], arr[-1]
    completed = np.count_nonzero((done >= t_lo) & (done <= t_hi))
    through = completed / (t_hi - t_lo) if t_hi > t_lo else float("inf")
    w = ttf[win]
    q = max(1, len(w) // 4)
    climb = float(np.median(w[-q:]) / max(np.median(w[:q]), 1e-9))
    out = {"requests": n, "p50_ttf_s": float(np.percentile(ttf, 50)),
           "p99_ttf_s": float(np.percentile(ttf, 99)), "max_ttf_s": float(ttf.max()),
           "scenarios_per_s": n / span if span > 0 else float("inf"), "span_s": float(span),
           "offered_req_per_s": float(offered),
           "steady": {"requests": int(len(w)), "p50_ttf_s": float(np.percentile(w, 50)),
                      "p99_ttf_s": float(np.percentile(w, 99)),
                      "completions_per_s": float(through), "ttf_climb": climb,
                      "sustained": bool(through >= 0.9 * offered and climb < 2.0)}}
    return out
