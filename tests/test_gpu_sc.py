"""GPU: SC_OPTION = 1 coupling in the device loop, bit for bit.

A Cartesian child of the spherical layer 1, laid out by the reference's
SUBGRID_MATCHING (POS, CXY), runs on the device with EDGE_INTERP_SC,
NEWQ_SC (band forecasts written into the parent's level 2, the edges
blended from them) and JNZ_SC (parent elevation averaged over the wet child
cells of each parent cell); the oracle, pinned against the reference's own
NEWQ_SC and JNZ_SC in tests/test_sc_oracle.py, must give the same bits for
every plane of both layers, the maxima and the arrival field.  A fault
spread to the child goes through CELL_INTERP_SC.  A batched SC nest
matches its run alone.
"""
import numpy as np
import pytest

from gcomcot import DeviceRun
from helpers import BEACH, diff_report, sc_nest
from oracle import okada as OK
from oracle.bind import Oracle, Reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(), reason="needs the flang-built reference")]


# REL_TIME 1 puts JNZ_SC in the first sub-step, before any UPDATE has
# filled the child's ghost row and column
@pytest.mark.parametrize("child_gov,rel_size,rel_time", [(1, 3, 3), (0, 3, 2), (1, 1, 1)])
def test_sc_nest_matches_the_oracle(child_gov, rel_size, rel_time):
    sc, _ = sc_nest(laygov_child=child_gov, rel_size=rel_size, rel_time=rel_time)
    n = 40
    print(f"[gpu] SC nest child laygov {child_gov} REL_SIZE {rel_size} REL_TIME {rel_time}, "
          f"{n} steps", flush=True)
    dev = DeviceRun(sc)
    dev.run(n // 2 + 1)
    dev.run(n - n // 2 - 1, graph=False)
    orc = Oracle(sc, 32, pow_mode=1)
    orc.step(n)
    for L in sc.layers:
        g, o = dev.state(L.id), orc.state(L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(g[k], o[k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
        gm, om = dev.maxima(L.id), orc.maxima(L.id)
        assert diff_report(gm["zmax_wet"], om["zmax_wet"])[0] == 0
        assert np.array_equal(gm["arrival"], om["arrival"])
    dev.close()


def test_fault_spreads_to_an_sc_child():
    sc, _ = sc_nest()
    R, C = sc.layers
    from gcomcot import scenarios as S
    x, y = S.layer_coords(R)
    f = OK.Fault(10000, 100000, 50000, 5, 30, 15, 90, x0=float(x[40]), y0=float(y[36]))
    print("[gpu] fault on an SC nest", flush=True)
    dev = DeviceRun(sc)
    dev.set_coords([1])
    dev.fault(f.params(), f.switch)
    d_root, d_child = dev.deform(1), dev.deform(2)
    dev.close()
    want = OK.cell_interp_sc(d_root, C.sc_pos, C.sc_cxy, C.nx, C.ny)
    nb, mx = diff_report(d_child, want)
    assert nb == 0, f"CELL_INTERP_SC: {nb} cells differ (max {mx})"
    assert np.abs(d_child).max() > 0.01


def test_batched_sc_nest_equals_its_single_run():
    """The batch engine's SC edge and JNZ groups, beside a traditional nest and
    a single layer, give the SC nest the bits of its run alone."""
    from gcomcot import scenarios as S
    from gcomcot.batch import Batch
    sc, _ = sc_nest()
    n = 24
    print(f"[gpu] batched SC nest, {n} steps", flush=True)
    dev = DeviceRun(sc)
    dev.run(n)
    alone = {L.id: dev.state(L.id) for L in sc.layers}
    dev.close()
    b = Batch()
    child = dict(laycord=0, laygov=1, footprint=(20, 24), rel_size=3, rel_time=3, corner=(30, 26))
    b.add(S.custom(laycord=0, laygov=1, nx=96, ny=80, dx=2.0, child=child, **BEACH), n)
    sid = b.add(sc, n)
    b.add(S.custom(laycord=1, laygov=1, nx=64, ny=56, **BEACH), n + 5)
    b.run(n + 5)
    for L in sc.layers:
        got = b.state(sid, L.id)
        for k in ("z1", "m1", "n1"):
            nb, mx = diff_report(got[k], alone[L.id][k])
            assert nb == 0, f"layer {L.id} {k}: {nb} cells differ (max {mx})"
    b.close()


def _planes(dev, sc):
    return {L.id: dev.state(L.id) for L in sc.layers}


def _same_planes(a, b, tag, keys=("z1", "m1", "n1")):
    for lid in a:
        for k in keys:
            nb, mx = diff_report(a[lid][k], b[lid][k])
            assert nb == 0, f"{tag} layer {lid} {k}: {nb} cells differ (max {mx})"


def test_reference_call_sequence_on_an_sc_nest_equals_the_device_loop():
    """MASS_LAUNCH, ALL_GRID_LAUNCH, MOMT_LAUNCH, MAX_AMP and the level copy
    per root step (comcot.f90:362-439) drive the SC kernels as the loop does."""
    sc, _ = sc_nest(rel_time=2)
    n = 12
    print(f"[gpu] SC nest through the drop-in calls, {n} steps", flush=True)
    dev = DeviceRun(sc)
    for _ in range(n):
        dev.dropin_step()
    got = _planes(dev, sc)
    dev.close()
    dev = DeviceRun(sc)
    dev.run(n)
    _same_planes(got, _planes(dev, sc), "drop-in vs loop")
    dev.close()


def test_sc_nest_checkpoint_continues_to_the_same_bits(tmp_path):
    sc, _ = sc_nest()
    ck = tmp_path / "sc.ckpt"
    print("[gpu] SC nest checkpoint after 13 steps, 10 more", flush=True)
    dev = DeviceRun(sc)
    dev.run(13)
    dev.save(ck)
    dev.run(10)
    want = _planes(dev, sc)
    dev.close()
    dev = DeviceRun(sc)
    dev.load(ck)
    dev.run(10)
    _same_planes(_planes(dev, sc), want, "restored", keys=("z1", "z2", "m1", "m2", "n1", "n2"))
    dev.close()


def test_sc_layer_errors_are_reported():
    import ctypes
    from gcomcot import lib
    sc, _ = sc_nest()
    R, C = sc.layers
    print("[gpu] SC layer errors", flush=True)

    def add(spec, pos=None, corners=None, parent=1):
        L = lib()
        pa = np.ascontiguousarray(spec.sc_pos if pos is None else pos, np.int32)
        ca = np.ascontiguousarray(spec.sc_cxy, np.float32)
        cr = np.array(spec.corners if corners is None else corners, np.int32)
        h = np.ascontiguousarray(spec.h, np.float32)
        z = np.ascontiguousarray(spec.z0, np.float32)
        I = lambda v: ctypes.byref(ctypes.c_int(v))
        F = lambda v: ctypes.byref(ctypes.c_float(v))
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        return L.gcomcot_init_layer_sc_(I(spec.id), I(parent), I(2), I(spec.laygov),
                                        I(spec.rel_size), I(spec.rel_time), P(cr),
                                        I(spec.fric_switch), F(spec.fric_coef), F(spec.dt),
                                        F(spec.dx), F(spec.dy), P(h), P(z), I(spec.nx),
                                        I(spec.ny), P(pa), P(ca))

    from gcomcot import scenarios as S
    from gcomcot._lib import last_error

    def refused(code, **kw):
        rc = add(C, **kw)
        got, msg = last_error()          # read (and clear) the recorded error
        assert rc == code and got == code, (rc, got, msg)

    dev = DeviceRun(S.Scenario("root-only", [R], 10))
    bad = np.array(C.sc_pos, np.int32).copy()
    bad[0] = C.corners[0] - 1            # a POS left of the extent
    refused(82, pos=bad)
    nx, ny = ctypes.c_int(0), ctypes.c_int(0)   # the refused layer is not left in the tree
    assert lib().gcomcot_layer_dims_(ctypes.byref(ctypes.c_int(C.id)), ctypes.byref(nx),
                                     ctypes.byref(ny)) != 0
    far = list(C.corners)
    far[1] = R.nx - 2                    # extent reaching the parent's last cells
    refused(5, corners=far)
    assert add(C) == 0                   # the good layer still goes in afterwards
    assert last_error()[0] == 0
    dev.close()
    cart = S.custom(laycord=1, laygov=0, nx=96, ny=80)
    dev = DeviceRun(cart)
    refused(80)                          # SC needs a spherical layer 1
    dev.close()


def test_batch_refuses_bad_sc_geometry_and_stays_usable():
    import copy
    from gcomcot._lib import GcomcotError
    from gcomcot.batch import Batch
    sc, _ = sc_nest()
    bad = copy.deepcopy(sc)
    pos = np.array(bad.layers[1].sc_pos, np.int32).copy()
    pos[0] = bad.layers[1].corners[0] - 1      # a POS left of the extent
    bad.layers[1].sc_pos = pos
    print("[gpu] batch with a refused SC layer", flush=True)
    b = Batch()
    with pytest.raises(GcomcotError, match="outside its extent"):
        b.add(bad, 4)
    sid = b.add(sc, 4)
    b.run(4)
    assert b.progress(sid) == (4, 4)
    b.close()
