"""Shared scenario lists and comparison helpers of the test suite."""
import numpy as np

from gcomcot import scenarios as S


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


def diff_report(a, b):
    """(cells whose float32 bit patterns differ, max |a-b| over finite cells).
    Two NaNs count as equal whatever their payloads."""
    a64, b64 = np.asarray(a, np.float64), np.asarray(b, np.float64)
    both_nan = np.isnan(a64) & np.isnan(b64)
    nb = int(np.count_nonzero((bits(a) != bits(b)) & ~both_nan))
    fin = np.isfinite(a64) & np.isfinite(b64)
    mx = float(np.max(np.abs(a64[fin] - b64[fin]))) if fin.any() else 0.0
    return nb, mx


BEACH = dict(embayments=2, beach_slope=3000.0, bay_floor=6.0, coast_frac=0.3, peak=3.0, cx=0.5)


def small_scenarios():
    """(name, scenario, steps) covering every layer kind and nest pairing."""
    out = [
        ("sph-linear", S.custom(laycord=0, laygov=0, nx=72, ny=64, cx=0.45), 150),
        ("cart-linear-fric-scm", S.custom(laycord=1, laygov=0, nx=72, ny=64, cx=0.45), 150),
        ("sph-nonlinear-shore", S.custom(laycord=0, laygov=1, nx=72, ny=64, **BEACH), 200),
        ("cart-nonlinear-shore", S.custom(laycord=1, laygov=1, nx=72, ny=64, **BEACH), 200),
        ("ragged-77x61", S.custom(laycord=0, laygov=1, nx=77, ny=61, **BEACH), 120),
    ]
    for clc, clg, R, RT in [(0, 0, 2, 2), (1, 1, 4, 6), (0, 1, 3, 5), (1, 0, 3, 3), (1, 1, 2, 3)]:
        sc = S.custom(nx=72, ny=64, laygov=0, child=dict(laycord=clc, laygov=clg,
                                                         footprint=(20, 36), rel_size=R,
                                                         rel_time=RT, corner=(8, 14)), **BEACH)
        out.append((f"nest-c{clc}{clg}-R{R}T{RT}", sc, 80))
    return out
