"""GPU: a layer whose planes exceed 2^31 bytes, against a small layer.

A Cartesian linear layer of 24576 x 24576 cells (2.4 GB per plane) and one
of 512 x 512 share depth, spacing and time step, and carry the same
Gaussian hump at the same offsets from a cell, placed near the far corner
of the large layer (plane offsets beyond 2^31 bytes).  The scheme reaches
at most two cells per step, so after n steps every cell within 256 - 2n of
the hump sees neither layer's edges, and holds the same bits in both:
indexing, tiling and launch shapes at the largest sizes change nothing.
"""
import ctypes

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot import scenarios as S
from gcomcot import lib

pytestmark = pytest.mark.gpu

STEPS = 60
FILL = 40          # hump radius in cells; zero beyond
SIGMA = 8.0


def _layer(n, ci, cj, depth, dx, dt):
    h = np.full((n, n), depth, np.float32)
    z0 = np.zeros((n, n), np.float32)
    d = np.arange(-FILL, FILL + 1)
    dj, di = np.meshgrid(d, d, indexing="ij")
    hump = (np.float32(1.5) * np.exp(-(di * di + dj * dj) / (2.0 * SIGMA * SIGMA))).astype(np.float32)
    z0[cj - 1 - FILL:cj + FILL, ci - 1 - FILL:ci + FILL] = hump
    spec = S.LayerSpec(id=1, parent=-1, level=1, laycord=1, laygov=0, nx=n, ny=n, dx=dx, dy=dx,
                       dt=dt, h=h, z0=z0, fric_switch=1, fric_coef=0.0)
    return S.Scenario(f"cartesian-linear-{n}x{n}", [spec], STEPS)


def _planes(n):
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    z = np.empty((n, n), np.float32)
    m = np.empty((n, n), np.float32)
    q = np.empty((n, n), np.float32)
    one = ctypes.byref(ctypes.c_int(1))
    lib().cuda_getz1_(p(z), one)
    lib().cuda_getmn1_(p(m), p(q), one)
    return z, m, q


def test_large_layer_equals_a_small_one_inside_the_domain_of_dependence():
    depth, dx = np.float32(4000.0), 1000.0
    dt = S.cr_dt(1, 0, np.full((4, 4), depth, np.float32), dx, dx, 0.0, 0.0)
    big_n, small_n = 24576, 512
    bi, bj = big_n - 400, big_n - 400          # hump cell of the large layer
    si, sj = small_n // 2, small_n // 2        # and of the small one
    print(f"[gpu] large layer {big_n}x{big_n}, {STEPS} steps", flush=True)
    dev = DeviceRun(_layer(small_n, si, sj, depth, dx, dt))
    dev.run(STEPS)
    small = _planes(small_n)
    dev.close()
    dev = DeviceRun(_layer(big_n, bi, bj, depth, dx, dt))
    dev.run(STEPS)
    big = _planes(big_n)
    dev.close()
    r = small_n // 2 - 2 * STEPS - 8
    for name, a, b in zip(("z", "m", "n"), small, big):
        wa = a[sj - 1 - r:sj + r, si - 1 - r:si + r]
        wb = b[bj - 1 - r:bj + r, bi - 1 - r:bi + r]
        assert np.abs(wa).max() > 1e-3, name
        assert np.array_equal(wa.view(np.int32), wb.view(np.int32)), \
            f"{name}: {np.count_nonzero(wa.view(np.int32) != wb.view(np.int32))} cells differ"
