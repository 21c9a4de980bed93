"""Branch coverage of the golden fixtures that pin CONMOME(_S).

The oracle counts, per flux point, which case of the reference's momentum
IF chains it took (moment.f90:280-780, 783-1268) and whether WAVE_BREAKING
(:1272-1384) damped the convection terms.  The golden fixtures compare the
oracle with the flang-built reference bit for bit (test_oracle_golden.py);
this file replays them and requires that, together, they reach every
branch, so that a pin of a branch rests on a fixture that takes it:
  * the finite damping COEF = EXP(-REF) at wet cells, with REF below and
    above 1 (DZDT between and beyond 0.08 and 0.65 sqrt(g D));
  * the four along-flux upwind cases and the six cross-term cases;
  * the DF clamp, both V_LIMIT clamps, friction on and off.
"""
import pytest

import fixtures
from oracle.bind import BRANCHES, Oracle

MOMENT_LOCAL = ["local_moment_sph_brk", "local_moment_cart_brk", "local_moment_cart_brk_nofric"]
SURF_RUNS = ["surf_sph_nonlinear", "surf_cart_nonlinear", "surf_cart_nonlinear_nofric"]


def _replay(name):
    f = fixtures.load(name)
    sc = fixtures.scenario(f)
    o = Oracle(sc, 32, pow_mode=0)
    o.branch_counts()
    if "op" in f:
        o.set_state(1, {k: f["in_" + k] for k in ("z1", "z2", "m1", "m2", "n1", "n2", "dz1",
                                                   "dz2")})
        o.call(int(f["op"][0]), 1, 0)
    else:
        o.step(sc.steps)
    c = o.branch_counts()
    print(name, {k: v for k, v in c.items() if v != (0, 0)})
    return {k: x + y for k, (x, y) in c.items()}


@pytest.mark.parametrize("name", MOMENT_LOCAL)
def test_step_local_fixture_breaks(name, oracle_built):
    c = _replay(name)
    assert c["brk_finite"] >= 1000, c
    assert c["brk_over"] >= 300, c       # REF > 1
    assert c["brk_none"] >= 100, c       # DZDT <= DZDT_F
    assert c["brk_finite"] - c["brk_over"] >= 500, c  # 0 < REF <= 1


@pytest.mark.parametrize("name", SURF_RUNS)
def test_surf_run_fixture_breaks(name, oracle_built):
    c = _replay(name)
    assert c["brk_finite"] >= 50, c
    assert c["dd_edge_a"] + c["dd_edge_b"] >= 1000, c
    assert c["df_clamp"] >= 10, c


def test_fixtures_reach_every_branch(oracle_built):
    total = {k: 0 for k in BRANCHES}
    for name in fixtures.names():
        f = fixtures.load(name)
        sc = fixtures.scenario(f)
        if all(L.laygov == 0 for L in sc.layers):
            continue
        if "op" in f and int(f["op"][0]) != 3:
            continue
        for k, v in _replay(name).items():
            total[k] += v
    missing = [k for k, v in total.items() if v == 0]
    assert not missing, f"no golden fixture takes {missing}: {total}"
    # friction both ways, and the nofric fixtures are frictionless throughout
    assert _replay("local_moment_cart_brk_nofric")["fric_on"] == 0
