"""Shared scenario lists and comparison helpers of the test suite."""
import math

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


def compare_device_oracle(sc, dev, orc, tag):
    """Every plane, the wet maximum and the arrival step of every layer of a
    DeviceRun against the oracle, bit for bit (and the interface flux cache
    of a nest)."""
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"{tag} layer {L.id} {k}: {nb} cells differ (max {mx})"
        gm, om = dev.maxima(L.id), orc.maxima(L.id)
        nb, mx = diff_report(gm["zmax_wet"], om["zmax_wet"])
        assert nb == 0, f"{tag} layer {L.id} zmax: {nb} cells differ (max {mx})"
        assert np.array_equal(gm["arrival"], om["arrival"]), f"{tag} layer {L.id} arrival differs"
    if len(sc.layers) > 1:
        gx, gy = dev.flux_cache(1)
        ox, oy = orc.flux(1)
        assert np.array_equal(bits(gx), bits(ox)) and np.array_equal(bits(gy), bits(oy)), \
            f"{tag} interface flux cache differs"


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


def surf_zone(laycord, nx=64, ny=56, dx_m=9.0, slope=20.0, hmax=6.0, peak=2.0, sig=2.2,
              fric_coef=0.013, cx=0.62, cy=0.5):
    """A nonlinear layer on a surf-zone scale: cells of ~dx_m metres (spherical
    layers take dx_m/1852 arc minutes), a 1:slope beach along the west side
    whose shoreline meanders by 3 cells, flat at hmax offshore, and a narrow
    elliptical hump of `peak` metres (sigma `sig` cells) in the interior.  On
    such a grid a front rises faster than 0.08*sqrt(g*D) per step, so the
    finite branch of WAVE_BREAKING (moment.f90:1319-1335) is taken, and the
    run-up reaches the V_LIMIT clamp (moment.f90:1055-1057); fric_coef 0
    switches friction off (FRIC_SWITCH = 1)."""
    dxa = float(np.float32(dx_m / 1852.0))
    sc = S.custom(laycord=laycord, laygov=1, nx=nx, ny=ny, dx=dxa, fric_coef=fric_coef)
    L = sc.layers[0]
    I, J = np.meshgrid(np.arange(1, nx + 1, dtype=np.float64),
                       np.arange(1, ny + 1, dtype=np.float64))
    xs = 0.25 * nx + 3.0 * np.sin(2.0 * math.pi * J / ny)
    h = np.minimum((I - xs) * dx_m / slope, hmax).astype(np.float32)
    r2 = ((I - cx * nx) ** 2 + ((J - cy * ny) * 1.3) ** 2) / sig ** 2
    z0 = np.where(h > 0, peak * np.exp(-0.5 * r2), 0.0).astype(np.float32)
    L.h, L.z0 = h, z0
    if laycord == 0:
        L.dt = S.cr_dt(0, 1, h, L.dx, L.dy, float(L.y[0]), float(L.y[-1]))
    else:
        L.dt = S.cr_dt(1, 1, h, L.dx, L.dy, 0.0, 0.0)
    sc.name = f"surf-{laycord}-{nx}x{ny}" + ("" if fric_coef > 0 else "-nofric")
    return sc


def breaking_state(rng, st, h, dt, frac=(0.0, 1.0), noise=0.5, thin=0.04):
    """A seeded level-1/level-2 state for one momentum call that puts every
    wet cell's |Z2 - Z1|/DT at u*sqrt(g*DZ2), u uniform in frac: Z2/DZ2 are
    kept (mass-consistent), flux noise of `noise` m^2/s is added, and Z1 is
    moved; WAVE_BREAKING is the only reader of Z1 in CONMOME(_S).  A `thin`
    fraction of the wet cells first gets a total depth in [1E-5, 2.5E-5) m
    (Z2 = D - H, DZ2 = Z2 + H in REAL, as CONMASS forms it), so that flux
    depths of two such neighbours fall below GX (moment.f90:996-1027)."""
    out = {k: np.array(v, np.float32, copy=True) for k, v in st.items()}
    for k in ("m1", "n1"):
        out[k] = (out[k] + rng.normal(0.0, noise, out[k].shape)).astype(np.float32)
    h = np.asarray(h, np.float32)
    pick = (out["dz2"] > 0) & (rng.random(h.shape) < thin)
    d = rng.uniform(1.0e-5, 2.5e-5, h.shape).astype(np.float32)
    z2t = (d - h).astype(np.float32)
    dzt = (z2t + h).astype(np.float32)
    ok = pick & (dzt >= np.float32(1.0e-5))
    out["z2"] = np.where(ok, z2t, out["z2"]).astype(np.float32)
    out["dz2"] = np.where(ok, dzt, out["dz2"]).astype(np.float32)
    dz2 = out["dz2"].astype(np.float64)
    u = rng.uniform(frac[0], frac[1], dz2.shape)
    sgn = np.where(rng.random(dz2.shape) < 0.5, -1.0, 1.0)
    dz = sgn * u * float(dt) * np.sqrt(9.807 * dz2)
    out["z1"] = np.where(dz2 > 0, out["z2"] - dz, out["z1"]).astype(np.float32)
    out["dz1"] = np.zeros_like(out["dz2"])
    return out


def sc_nest(laygov_child=1, rel_size=3, rel_time=3, box=(30, 50, 26, 46), root=None):
    """A Cartesian child of a spherical layer 1 with SC_OPTION = 1, laid out by
    the reference's own SUBGRID_MATCHING (flang build): returns the Scenario
    (child depth and elevation bilinear from the parent's through POS/CXY)
    and the Reference holding the same two layers."""
    from oracle.bind import Reference
    root = root or S.custom(laycord=0, laygov=0, nx=96, ny=80, dx=2.0, **BEACH)
    R = root.layers[0]
    x, y = S.layer_coords(R)
    ref = Reference(root)
    ref.set_coords(1, x, y)
    ext = [x[box[0]], x[box[1]], y[box[2]], y[box[3]]]
    dt = float(np.float32(np.float32(R.dt) / np.float32(rel_time)))
    fs, fc = (0, 0.013) if laygov_child else (1, 0.0)
    g = ref.sc_child(laygov_child, rel_size, rel_time, ext, dt, fs, fc)
    nx, ny = g["nx"], g["ny"]
    pos = g["pos"].reshape(2, ny, nx)
    cxy = g["cxy"].reshape(4, ny, nx)

    def spread(plane):
        P = np.asarray(plane, np.float32)
        ki, kj = pos[0] - 1, pos[1] - 1
        return (((P[kj, ki] * cxy[0] + P[kj, ki + 1] * cxy[1]) + P[kj + 1, ki] * cxy[2])
                + P[kj + 1, ki + 1] * cxy[3]).astype(np.float32)

    h, z0 = spread(R.h), spread(R.z0)
    ref.set_hz(g["id"], h, z0)
    child = S.LayerSpec(id=g["id"], parent=1, level=2, laycord=1, laygov=laygov_child, nx=nx,
                        ny=ny, dx=g["dx"], dy=g["dy"], dt=dt, h=h, z0=z0, fric_switch=fs,
                        fric_coef=fc, rel_size=rel_size, rel_time=rel_time,
                        corners=g["corners"], sc_pos=g["pos"], sc_cxy=g["cxy"])
    return S.Scenario("sc-nest", [R, child], 40), ref
