"""The serving summary's steady-state test on synthetic request streams
(host only): a stream served faster than it arrives is sustained, one whose
queue grows is not."""
import numpy as np

from gcomcot.serve import summarize


class _R:
    def __init__(self, arrival, done):
        self.arrival, self.done = arrival, done


def _stream(rate, service, n=200):
    arr = np.arange(n) / rate
    done, free = [], 0.0
    for a in arr:  # one server, first come first served
        free = max(free, a) + service
        done.append(free)
    return [_R(a, d) for a, d in zip(arr, done)]


def test_sustained_stream():
    s = summarize(_stream(rate=10.0, service=0.05))
    assert s["steady"]["sustained"]
    assert abs(s["offered_req_per_s"] - 10.0) < 1e-9
    assert abs(s["p50_ttf_s"] - 0.05) < 1e-9 and abs(s["steady"]["p99_ttf_s"] - 0.05) < 1e-9


def test_overloaded_stream():
    s = summarize(_stream(rate=10.0, service=0.2))
    assert not s["steady"]["sustained"]
    assert s["steady"]["ttf_climb"] > 2.0
    assert s["steady"]["completions_per_s"] < 0.9 * 10.0
    assert s["max_ttf_s"] > 10.0
