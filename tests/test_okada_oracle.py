"""The fault-source oracle (oracle/okada.py), pinned by the reference.

DEFORM_OKADA evaluates a sum of near-cancelling terms in REAL: its own
result moves by up to ~1e-3 m when a single fault parameter, or the cell
coordinates, move by one float ulp.  The FP64 restatement must agree with
the flang-built reference to within twice that sensitivity, measured here
on the reference itself, for
spherical and Cartesian layers, both epicentre conventions and thrust,
oblique and strike-slip faults.  A dropped term, a wrong sign or a
swapped angle moves the field by O(1) of its size and fails.

CELL_INTERP is REAL arithmetic and must match bit for bit: the
reference's own GET_FLOOR_DEFORM on a nest gives the parent and child
deformation; the oracle spreads the parent's field and must reproduce the
child's exactly, and UPDATE_BATH / the elevation update must be exactly
H - D and Z + D.
"""
import numpy as np
import pytest

from gcomcot import scenarios as S
from oracle import okada as OK
from oracle.bind import Reference

needs_ref = pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")


def _faults(xc, yc):
    return [OK.Fault(10000, 100000, 50000, 5, 30, 15, 90, x0=xc, y0=yc),
            OK.Fault(5000, 60000, 30000, 3, 200, 45, 30, x0=xc, y0=yc, switch=9),
            OK.Fault(20000, 150000, 60000, 8, 10, 80, 0, x0=xc, y0=yc)]


def _sensitivity(ref, params, fswitch, base, coords=None):
    """Largest change of the reference's DEFORM when one fault parameter, or
    the layer's longitudes or latitudes, move by one float ulp either way."""
    worst = 0.0
    up, down = np.float32(np.inf), np.float32(-np.inf)
    for k in range(len(params)):
        for direction in (up, down):
            q = params.copy()
            q[k] = np.nextafter(q[k], direction)
            worst = max(worst, float(np.abs(ref.okada(q, fswitch).astype(np.float64) - base).max()))
    if coords is not None:
        x, y = (np.asarray(c, np.float32) for c in coords)
        for xs, ys in ((np.nextafter(x, up), y), (np.nextafter(x, down), y),
                       (x, np.nextafter(y, up)), (x, np.nextafter(y, down))):
            ref.set_coords(1, xs, ys)
            worst = max(worst, float(np.abs(ref.okada(params, fswitch).astype(np.float64)
                                            - base).max()))
        ref.set_coords(1, x, y)
    return worst


def _check(ref, f, want, coords=None):
    p = f.params()
    got = ref.okada(p, f.switch).astype(np.float64)
    tol = 2.0 * _sensitivity(ref, p, f.switch, got, coords)
    diff = np.abs(got - want)
    assert np.abs(want).max() > 0.01
    assert diff.max() <= tol, f"max diff {diff.max():.3e} > {tol:.3e}"
    # the field, not just its peak: the mean difference is far below the peak one
    assert diff.mean() <= 0.25 * tol
    # cells zeroed by the GX threshold agree except right at the threshold
    zr, zo = got == 0, want == 0
    assert np.all(np.abs(np.where(zr != zo, got + want, 0)) <= 4 * OK.GX)


@needs_ref
@pytest.mark.parametrize("dx", [2.0, 0.5])
def test_spherical_okada_matches_the_reference(dx):
    sc = S.custom(laycord=0, laygov=0, nx=96, ny=80, dx=dx)
    L = sc.layers[0]
    x, y = S.layer_coords(L)
    ref = Reference(sc)
    ref.set_coords(1, x, y)
    for f in _faults(float(x[48]), float(y[40])):
        _check(ref, f, OK.deform_okada(f, L.nx, L.ny, 0, x, y).astype(np.float64), (x, y))


@needs_ref
def test_cartesian_okada_matches_the_reference():
    sc = S.custom(laycord=1, laygov=0, nx=96, ny=80)
    L = sc.layers[0]
    ref = Reference(sc)
    for f in _faults(140.4, 10.3):
        f.xo, f.yo = 140.0, 10.0
        _check(ref, f, OK.deform_okada(f, L.nx, L.ny, 1, dx=L.dx, dy=L.dy).astype(np.float64))


@pytest.mark.parametrize("rake", [0.0, 60.0])
def test_okada_is_linear_in_the_slip_and_zero_without_it(rake):
    x = np.linspace(139.0, 141.0, 40, dtype=np.float32)
    y = np.linspace(9.0, 11.0, 36, dtype=np.float32)
    f = OK.Fault(10000, 100000, 50000, 2, 30, 15, rake, x0=140.0, y0=10.0)
    d1 = OK.deform_okada(f, 40, 36, 0, x, y)
    f.d = 4
    d2 = OK.deform_okada(f, 40, 36, 0, x, y)
    big = np.abs(d1) > 1e-3
    assert big.sum() > 100
    if rake == 0.0:  # one term: slip 2 -> 4 doubles it exactly
        np.testing.assert_array_equal(d2[big], 2 * d1[big])
    else:            # two terms, each cut at GX separately
        np.testing.assert_allclose(d2[big], 2 * d1[big], rtol=0, atol=2 * OK.GX)
    f.d = 0
    assert not OK.deform_okada(f, 40, 36, 0, x, y).any()


def test_projection_of_the_epicentre_is_the_origin_and_north_is_up():
    x, y = OK.stereo_projection(np.array([140.0, 140.0, 140.5]), np.array([10.0, 10.5, 10.0]),
                                140.0, 10.0)
    assert abs(x[0]) < 1e-6 and abs(y[0]) < 1e-6
    assert abs(x[1]) < 1.0 and 55e3 < y[1] < 56e3        # half a degree north
    assert 54e3 < x[2] < 55e3 and abs(y[2]) < 500.0        # half a degree east at 10 N


@needs_ref
def test_cell_interp_and_bathymetry_update_match_the_reference():
    child = dict(laycord=0, laygov=0, footprint=(20, 24), rel_size=3, rel_time=3, corner=(30, 26))
    sc = S.custom(laycord=0, laygov=0, nx=96, ny=80, dx=2.0, child=child)
    ref = Reference(sc)
    for L in sc.layers:
        ref.set_coords(L.id, *S.layer_coords(L))
    before = {L.id: (ref.deform(L.id)[1].copy(), ref.state(L.id)["z1"].copy()) for L in sc.layers}
    x, y = S.layer_coords(sc.layers[0])
    f = _faults(float(x[45]), float(y[38]))[0]
    ref.floor_deform(f.params(), f.switch)
    d_root = ref.deform(1)[0]
    C = sc.layers[1]
    d_child = ref.deform(2)[0]
    mine = OK.cell_interp(d_root, (C.ny, C.nx), C.corners, C.rel_size)
    assert np.abs(d_child).max() > 0.1
    np.testing.assert_array_equal(mine.view(np.int32), d_child.view(np.int32))
    for L in sc.layers:
        d, h = ref.deform(L.id)
        h0, z0 = before[L.id]
        np.testing.assert_array_equal(h, h0 - d)
        np.testing.assert_array_equal(ref.state(L.id)["z1"], z0 + d)
