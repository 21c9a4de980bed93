"""The oracle against the reference Fortran's golden vectors, bit for bit.

tests/golden/*.npz were written by tests/golden/make_golden.py from the
flang-built reference (oracle/_ref).  Whole-run fixtures cover every layer
kind and nest pairing; step-local fixtures apply one routine to a seeded
random state, as the reference's DEBUG_CORE mode does.
"""
import numpy as np
import pytest

import fixtures
from helpers import bits
from oracle.bind import OP, Oracle

RUNS = [n for n in fixtures.names() if not n.startswith("local_")]
LOCAL = fixtures.names("local_")


@pytest.mark.parametrize("pow_mode", [0, 1], ids=["libm", "device-math"])
@pytest.mark.parametrize("name", RUNS)
def test_run_matches_reference(name, pow_mode, oracle_built):
    f = fixtures.load(name)
    sc = fixtures.scenario(f)
    o = Oracle(sc, 32, pow_mode=pow_mode)
    o.step(sc.steps)
    for L in sc.layers:
        st = o.state(L.id)
        for k in ("z1", "m1", "n1", "dz2"):
            want = f[f"out{L.id}_{k}"]
            assert np.array_equal(bits(st[k]), bits(want)), f"{name} layer {L.id} {k}"
    if "out1_xflux" in f:
        xf, yf = o.flux(1)
        assert np.array_equal(bits(xf), bits(f["out1_xflux"]))
        assert np.array_equal(bits(yf), bits(f["out1_yflux"]))


@pytest.mark.parametrize("name", LOCAL)
def test_routine_matches_reference(name, oracle_built):
    f = fixtures.load(name)
    sc = fixtures.scenario(f)
    op, lid, t = (int(v) for v in f["op"])
    o = Oracle(sc, 32, pow_mode=0)
    o.set_state(lid, {k: f["in_" + k] for k in ("z1", "z2", "m1", "m2", "n1", "n2", "dz1", "dz2")})
    o.call(op, lid, t)
    st = o.state(lid)
    for k in ("z2", "m2", "n2", "dz2"):
        assert np.array_equal(bits(st[k]), bits(f["out_" + k])), f"{name} {k}"


def test_fixtures_present():
    assert len(RUNS) >= 7 and len(LOCAL) >= 5
