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
