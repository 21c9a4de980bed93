"""GPU: the nonlinear momentum branches that only surf-zone states reach.

WAVE_BREAKING damps the convection terms by EXP(-REF) where a wet cell's
surface moves faster than 0.08*sqrt(g*D) per second (moment.f90:1300-1335);
basin-scale grids never get there.  These cases do, on ~9 m cells: whole
runs of a narrow hump onto a 1:20 beach (spherical and Cartesian, with and
without friction), and three steps from a seeded state with large flux
noise and planted pairs of cells whose total depth lies in [1E-5, 2.5E-5)
m.  Each test first counts, on the oracle, how many flux points took each
branch of CONMOME(_S) and requires the ones it is there for, so the
coverage cannot vanish unnoticed; then the device must give the oracle's
bits (device-math pow/exp) for every plane, the wet maximum and arrival.
Tiles of the 160 x 128 layers lie more than SPAN cells inside the grid
around the shoreline (I ~ 40), so the interior instantiation of the step
kernel takes these branches too.
"""
import ctypes

import numpy as np
import pytest

from gcomcot import DeviceRun
from gcomcot.batch import Batch
from helpers import compare_device_oracle, diff_report, surf_zone
from oracle.bind import Oracle

pytestmark = pytest.mark.gpu

SURF = [
    ("surf-sph", surf_zone(0, nx=160, ny=128, cx=0.45), 180),
    ("surf-cart", surf_zone(1, nx=160, ny=128, cx=0.45), 260),
    ("surf-cart-nofric", surf_zone(1, nx=160, ny=128, cx=0.45, fric_coef=0.0), 260),
]


def _counts(orc):
    return {k: x + y for k, (x, y) in orc.branch_counts().items()}


def _require(c, tag, **floors):
    for k, v in floors.items():
        assert c[k] >= v, f"{tag}: branch {k} taken {c[k]} times, need {v}"


def _oracle_run(sc, steps):
    orc = Oracle(sc, 32, pow_mode=1)
    orc.branch_counts()
    orc.step(steps)
    c = _counts(orc)
    print(f"[gpu] {sc.name}: oracle branches {c}", flush=True)
    return orc, c


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
@pytest.mark.parametrize("name,sc,steps", SURF, ids=[c[0] for c in SURF])
def test_surf_zone_run_matches_oracle(name, sc, steps, graph):
    print(f"[gpu] {name} graph={graph} steps={steps}", flush=True)
    orc, c = _oracle_run(sc, steps)
    _require(c, name, brk_finite=40, dd_edge_a=100, dd_edge_b=100, df_clamp=10,
             up_neg_deg=1, up_pos_deg=1, cross_neg_dz=1, cross_pos_dz=1)
    assert c["vlim_hi"] + c["vlim_lo"] >= 1, f"{name}: V_LIMIT never clamps"
    _require(c, name, **({"fric_off": 1} if sc.layers[0].fric_switch == 1 else {"fric_on": 1}))
    dev = DeviceRun(sc)
    dev.run(steps, graph=graph)
    compare_device_oracle(sc, dev, orc, name)
    dev.close()


def test_surf_zone_every_tile_shape():
    from gcomcot._lib import ip, lib
    name, sc, steps = SURF[2]
    n = ctypes.c_int(0)
    lib().gcomcot_tile_variants_(ctypes.byref(n), None, None)
    orc, c = _oracle_run(sc, steps)
    _require(c, name, brk_finite=40)
    try:
        for v in range(n.value):
            print(f"[gpu] {name} tile variant {v}", flush=True)
            assert lib().gcomcot_set_tile_variant_(ip(v)) == 0
            dev = DeviceRun(sc)
            dev.run(steps, graph=False)
            compare_device_oracle(sc, dev, orc, f"{name} variant {v}")
            dev.close()
    finally:
        lib().gcomcot_set_tile_variant_(ip(0))


def _noisy_state(sc, seed, settle=120, noise=4.0, pairs=300):
    """Level 1 of the surf zone after `settle` steps, with N(0, noise) added to
    the fluxes, level 2 equal to level 1, and `pairs` horizontally and
    vertically adjacent cell pairs made 'thin' (total depth in [1E-5,
    2.5E-5) m, zero flux on every face of the pair, so continuity keeps Z
    exactly and DZ2 = Z + H in REAL)."""
    rng = np.random.default_rng(seed)
    base = Oracle(sc, 32, pow_mode=1)
    base.step(settle)
    st = base.state(1)
    L = sc.layers[0]
    h = np.asarray(L.h, np.float32)
    z, m, n = (np.array(st[k], np.float32) for k in ("z1", "m1", "n1"))
    m = (m + rng.normal(0.0, noise, m.shape)).astype(np.float32)
    n = (n + rng.normal(0.0, noise, n.shape)).astype(np.float32)
    # CONMOME(_S) never writes the fluxes of the last column and row of the
    # outermost layer (moment.f90:377-387, 873-883), so a run holds 0 there
    for f in (m, n):
        f[:, -1] = 0.0
        f[-1, :] = 0.0
    ny, nx = h.shape
    planted = 0
    for _ in range(pairs * 4):
        if planted >= pairs:
            break
        i, j = int(rng.integers(14, nx - 14)), int(rng.integers(14, ny - 14))
        cells = [(j, i), (j, i + 1)] if rng.random() < 0.5 else [(j, i), (j + 1, i)]
        if not all(h[c] > 0.05 for c in cells):
            continue
        for (jj, ii) in cells:
            d = np.float32(rng.uniform(1.0e-5, 2.5e-5))
            zt = np.float32(d - h[jj, ii])
            if np.float32(zt + h[jj, ii]) < np.float32(1.0e-5):
                zt = np.float32(zt + np.float32(1.0e-5))
            z[jj, ii] = zt
            m[jj, ii] = 0
            m[jj, ii - 1] = 0
            n[jj, ii] = 0
            n[jj - 1, ii] = 0
        planted += 1
    zero = np.zeros_like(z)
    # level 2 as the level copy leaves it (comcot.f90:475-512): OPEN reads
    # the level-2 fluxes before MOMENT rewrites them (boundaries.f90:22-202)
    return dict(z1=z, m1=m, n1=n, z2=z.copy(), m2=m.copy(), n2=n.copy(), dz1=zero, dz2=zero)


@pytest.mark.parametrize("laycord,fric", [(0, 0.013), (1, 0.013), (1, 0.0)],
                         ids=["sph", "cart", "cart-nofric"])
def test_noisy_surf_state_matches_oracle(laycord, fric):
    sc = surf_zone(laycord, nx=160, ny=128, cx=0.45, fric_coef=fric)
    st = _noisy_state(sc, seed=31 + laycord)
    steps = 3
    orc = Oracle(sc, 32, pow_mode=1)
    orc.set_state(1, st)
    orc.branch_counts()
    orc.step(steps)
    c = _counts(orc)
    print(f"[gpu] noisy {sc.name}: oracle branches {c}", flush=True)
    _require(c, sc.name, brk_finite=2000, brk_over=100, brk_none=1000, brk_dry=1,
             dd_wet=1000, dd_edge_a=100, dd_edge_b=100, df_clamp=5,
             up_neg_deg=10, up_neg_full=10, up_pos_deg=10, up_pos_full=10,
             cross_neg_dz=10, cross_neg_full=10, cross_pos_dz=10, cross_pos_full=10,
             vlim_hi=10, vlim_lo=10)
    assert c["cross_neg_dp"] + c["cross_pos_dp"] >= 1, "no dry flux depth in the cross term"
    _require(c, sc.name, **({"fric_off": 1} if fric == 0.0 else {"fric_on": 1}))
    dev = DeviceRun(sc)
    dev.set_level1(1, z=st["z1"], m=st["m1"], n=st["n1"])
    dev.run(steps, graph=False)
    compare_device_oracle(sc, dev, orc, f"noisy {sc.name}")
    dev.close()


def test_batched_surf_zones_match_oracle():
    """The persistent batch kernel takes the same branches: three surf zones
    and their odd/even step counts in one batch, each against the oracle."""
    b = Batch()
    runs = [(name, sc, steps - k) for k, (name, sc, steps) in enumerate(SURF)]
    sids = [b.add(sc, n) for _, sc, n in runs]
    assert b.run(10 ** 6) == len(runs)
    for (name, sc, n), sid in zip(runs, sids):
        print(f"[gpu] batched {name} ({n} steps)", flush=True)
        orc, c = _oracle_run(sc, n)
        _require(c, name, brk_finite=40)
        g, o = b.state(sid, 1), orc.state(1)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"batched {name} {k}: {nb} cells differ (max {mx})"
        gm, om = b.maxima(sid, 1), orc.maxima(1)
        nb, mx = diff_report(gm["zmax_wet"], om["zmax_wet"])
        assert nb == 0, f"batched {name} zmax: {nb} differ"
        assert np.array_equal(gm["arrival"], om["arrival"]), f"batched {name} arrival"
    b.close()
