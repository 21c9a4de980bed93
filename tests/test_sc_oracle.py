"""SC_OPTION = 1 coupling: the oracle against the reference's own routines.

The reference lays out a Cartesian child of a spherical layer with its
SUBGRID_MATCHING (reverse UTM projection, POS and CXY).  On seeded random
states, the oracle's NEWQ_SC (band forecasts of the parent's fluxes and the
blended child edge fluxes, top edge with its C1*YN1) and JNZ_SC must equal
the flang-built reference bit for bit.  EDGE_INTERP_SC indexes its wet tests
with stale loop variables (undefined in the reference); the oracle tests the
boundary cell itself, and its interpolation is pinned by an affine parent
field, which bilinear weights must reproduce at the projected position.
"""
import numpy as np
import pytest

from helpers import diff_report, sc_nest
from oracle.bind import Oracle, Reference

needs_ref = pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")


def _random_state(rng, L, lo=0.0):
    st = {}
    for k in ("z", "m", "n", "dz"):
        for lev in "12":
            st[k + lev] = (rng.normal(lo, 1.0, (L.ny, L.nx))).astype(np.float32)
    return st


@needs_ref
@pytest.mark.parametrize("op,t", [(5, 2), (5, 3), (6, 0)])
def test_newq_sc_and_jnz_sc_match_the_reference(op, t):
    sc, ref = sc_nest()
    orc = Oracle(sc, 32, pow_mode=0)
    rng = np.random.default_rng(100 + 10 * op + t)
    for _ in range(2):
        for L in sc.layers:
            st = _random_state(rng, L)
            ref.set_state(L.id, st)
            orc.set_state(L.id, st)
        ref.call(op, 2, t)
        orc.call(op, 2, t)
        for L in sc.layers:
            a, b = orc.state(L.id), ref.state(L.id)
            for k in ("z1", "z2", "m1", "m2", "n1", "n2"):
                nb, mx = diff_report(a[k], b[k])
                assert nb == 0, f"op {op} t {t} layer {L.id} {k}: {nb} cells differ (max {mx})"


@needs_ref
def test_sc_geometry_from_the_reference():
    sc, _ = sc_nest()
    C = sc.layers[1]
    pos = C.sc_pos.reshape(2, C.ny, C.nx)
    cxy = C.sc_cxy.reshape(4, C.ny, C.nx)
    is_, ie, js, je = C.corners
    assert pos[0].min() == is_ and pos[0].max() == ie and pos[1].min() == js and pos[1].max() == je
    np.testing.assert_allclose(cxy.sum(axis=0), 1.0, atol=2e-6)
    assert (cxy >= 0).all()


@needs_ref
def test_edge_interp_sc_reproduces_an_affine_parent_field():
    sc, _ = sc_nest()
    R, C = sc.layers
    orc = Oracle(sc, 32, pow_mode=0)
    jj, ii = np.meshgrid(np.arange(1, R.ny + 1), np.arange(1, R.nx + 1), indexing="ij")
    fm = (0.5 + 0.01 * ii - 0.02 * jj).astype(np.float32)
    fn = (-0.3 + 0.015 * ii + 0.005 * jj).astype(np.float32)
    zero = np.zeros((R.ny, R.nx), np.float32)
    orc.set_state(1, dict(z1=zero, z2=zero, m1=fm, m2=fm, n1=fn, n2=fn, dz1=zero, dz2=zero))
    czero = np.zeros((C.ny, C.nx), np.float32)
    orc.set_state(2, dict(z1=czero, z2=czero, m1=czero, m2=czero, n1=czero, n2=czero,
                          dz1=czero, dz2=czero))
    orc.call(4, 2, 1)
    st = orc.state(2)
    pos = C.sc_pos.reshape(2, C.ny, C.nx).astype(np.float64)
    cxy = C.sc_cxy.reshape(4, C.ny, C.nx).astype(np.float64)
    fx = pos[0] + cxy[1] + cxy[3]     # KI + CX
    fy = pos[1] + cxy[2] + cxy[3]     # KJ + CY
    want_m = 0.5 + 0.01 * fx - 0.02 * fy
    want_n = -0.3 + 0.015 * fx + 0.005 * fy
    wet = C.h > 1e-5
    for (j, i) in [(slice(None), 0), (slice(None), -1)]:
        ok = wet[j, i]
        np.testing.assert_allclose(st["m1"][j, i][ok], want_m[j, i][ok], rtol=0, atol=2e-5)
    for (j, i) in [(0, slice(None)), (-1, slice(None))]:
        ok = wet[j, i]
        np.testing.assert_allclose(st["n1"][j, i][ok], want_n[j, i][ok], rtol=0, atol=2e-5)
    assert wet[:, 0].sum() > 10


@needs_ref
def test_cell_interp_sc_matches_the_reference():
    from gcomcot import scenarios as S
    from oracle import okada as OK
    sc, ref = sc_nest()
    R, C = sc.layers
    x, y = S.layer_coords(R)
    f = OK.Fault(10000, 100000, 50000, 5, 30, 15, 90, x0=float(x[40]), y0=float(y[36]))
    ref.floor_deform(f.params(), f.switch)
    d_root, d_child = ref.deform(1)[0], ref.deform(2)[0]
    mine = OK.cell_interp_sc(d_root, C.sc_pos, C.sc_cxy, C.nx, C.ny)
    assert np.abs(d_child).max() > 0.01
    np.testing.assert_array_equal(mine.view(np.int32), d_child.view(np.int32))


@needs_ref
@pytest.mark.parametrize("gov,rel_size,rel_time", [(1, 3, 3), (0, 3, 2), (1, 1, 1)])
def test_first_root_step_matches_the_reference(gov, rel_size, rel_time):
    """The whole ALL_GRID sequence of one root step (EDGE_INTERP_SC, NEWQ_SC,
    JNZ_SC over the child's ghost row and column, UPDATE) on the reference's
    own step.  Later steps part ways where EDGE_INTERP_SC's undefined wet
    tests (stale I, J = NY + 1) read memory outside the layer."""
    sc, ref = sc_nest(laygov_child=gov, rel_size=rel_size, rel_time=rel_time)
    orc = Oracle(sc, 32, pow_mode=0)
    ref.step(1)
    orc.step(1)
    for L in sc.layers:
        a, b = orc.state(L.id), ref.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(a[k], b[k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
