"""Write the golden fixtures of tests/golden/ from the reference Fortran.

Each fixture holds a small scenario's inputs (every array and scalar the
layers are built from) and the reference's outputs after a number of root
steps, computed by the flang-built reference (oracle/_ref, see
oracle/Makefile).  Step-local fixtures hold one routine applied to a
seeded random state, as the reference's DEBUG_CORE mode compares them.

Run in the build container (needs /root/reference and oracle/_ref):
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-comcot_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from gcomcot import scenarios as S  # noqa: E402
from helpers import breaking_state, surf_zone  # noqa: E402
from oracle import bind  # noqa: E402

BEACH = dict(embayments=2, beach_slope=3000.0, bay_floor=6.0, coast_frac=0.3, peak=3.0, cx=0.5)


def scenario_list():
    child = lambda lc, lg, R, RT: dict(laycord=lc, laygov=lg, footprint=(10, 14), rel_size=R,
                                       rel_time=RT, corner=(6, 8))
    return [
        ("sph_linear", S.custom(laycord=0, laygov=0, nx=48, ny=40, cx=0.45), 60),
        ("cart_linear_fric_scm", S.custom(laycord=1, laygov=0, nx=48, ny=40, cx=0.45), 60),
        ("sph_nonlinear_shore", S.custom(laycord=0, laygov=1, nx=48, ny=40, **BEACH), 80),
        ("cart_nonlinear_shore", S.custom(laycord=1, laygov=1, nx=48, ny=40, **BEACH), 80),
        ("nest_sph_sph_R2T2", S.custom(nx=48, ny=40, laygov=0, child=child(0, 0, 2, 2), **BEACH), 40),
        ("nest_sph_cartNL_R3T4", S.custom(nx=48, ny=40, laygov=0, child=child(1, 1, 3, 4), **BEACH),
         40),
        ("nest_sph_sphNL_R3T5", S.custom(nx=48, ny=40, laygov=0, child=child(0, 1, 3, 5), **BEACH),
         30),
        # surf-zone runs: finite WAVE_BREAKING and the V_LIMIT clamp
        ("surf_sph_nonlinear", surf_zone(0), 160),
        ("surf_cart_nonlinear", surf_zone(1), 160),
        ("surf_cart_nonlinear_nofric", surf_zone(1, fric_coef=0.0), 160),
    ]


def spec_arrays(sc):
    out = {}
    for L in sc.layers:
        p = f"L{L.id}_"
        out[p + "h"] = L.h
        out[p + "z0"] = L.z0
        if L.y is not None:
            out[p + "y"] = L.y
            out[p + "dely"] = L.dely
        out[p + "int"] = np.array([L.id, L.parent, L.level, L.laycord, L.laygov, L.nx, L.ny,
                                   L.fric_switch, L.rel_size, L.rel_time, *L.corners], np.int64)
        out[p + "flt"] = np.array([L.dx, L.dy, L.dt, L.fric_coef], np.float32)
    return out


def digest(arrays):
    h = hashlib.sha256()
    for k in sorted(arrays):
        h.update(k.encode())
        h.update(np.ascontiguousarray(arrays[k]).tobytes())
    return h.hexdigest()


def run_fixture(name, sc, steps):
    ref = bind.Reference(sc)
    ref.step(steps)
    arrays = spec_arrays(sc)
    for L in sc.layers:
        st = ref.state(L.id)
        for k in ("z1", "m1", "n1", "dz2"):
            arrays[f"out{L.id}_{k}"] = st[k]
    if len(sc.layers) > 1:
        xf, yf = ref.flux(1)
        arrays["out1_xflux"], arrays["out1_yflux"] = xf, yf
    arrays["steps"] = np.array([steps], np.int64)
    return arrays


def step_local(name, sc, op, seed, t=0, lid=1):
    """One routine on a seeded random state (DEBUG_CORE-style fixture)."""
    rng = np.random.default_rng(seed)
    ref = bind.Reference(sc)
    ref.step(3)  # a physical state to perturb
    st = {}
    base = ref.state(lid)
    L = sc.layers[lid - 1]
    for k in ("z", "m", "n"):
        a = base[k + "1"]
        noise = rng.normal(0.0, 0.05 if k == "z" else 2.0, a.shape).astype(np.float32)
        st[k + "1"] = (a + noise).astype(np.float32)
        st[k + "2"] = st[k + "1"].copy()   # OPEN reads level 2 = level 1 (after CHANGE)
    st["dz1"] = np.zeros_like(base["dz1"])
    st["dz2"] = np.maximum(st["z1"] + L.h, 0).astype(np.float32)
    st["dz2"][st["dz2"] < 1e-5] = 0
    if op == "MOMENT":  # momentum reads the new elevation: make it mass-consistent
        ref.set_state(lid, st)
        ref.call("MASS", lid)
        st = ref.state(lid)
    ref.set_state(lid, st)
    ref.call(op, lid, t)
    out = ref.state(lid)
    arrays = spec_arrays(sc)
    for k in ("z1", "z2", "m1", "m2", "n1", "n2", "dz1", "dz2"):
        arrays["in_" + k] = st[k]
        arrays["out_" + k] = out[k]
    arrays["op"] = np.array([bind.OP[op], lid, t], np.int64)
    return arrays


def step_local_breaking(sc, seed, settle):
    """CONMOME(_S) once on a state whose wet cells rise or fall at u times
    sqrt(g*D) per second, u uniform in [0, 1): below, between and above the
    0.08 and 0.65 thresholds of WAVE_BREAKING (moment.f90:1300-1335)."""
    rng = np.random.default_rng(seed)
    ref = bind.Reference(sc)
    ref.step(settle)
    st = ref.state(1)
    for k in ("z", "m", "n"):
        st[k + "2"] = st[k + "1"].copy()
    ref.set_state(1, st)
    ref.call("MASS", 1)  # level-2 elevation and total depth, consistent
    st = breaking_state(rng, ref.state(1), sc.layers[0].h, sc.layers[0].dt)
    ref.set_state(1, st)
    ref.call("MOMENT", 1)
    out = ref.state(1)
    arrays = spec_arrays(sc)
    for k in ("z1", "z2", "m1", "m2", "n1", "n2", "dz1", "dz2"):
        arrays["in_" + k] = st[k]
        arrays["out_" + k] = out[k]
    arrays["op"] = np.array([bind.OP["MOMENT"], 1, 0], np.int64)
    return arrays


def main():
    bind.build(ref=True)
    flang = subprocess.run(["/opt/rocm/lib/llvm/bin/flang", "--version"], capture_output=True,
                           text=True).stdout.splitlines()[0]
    manifest = {"generator": "tests/golden/make_golden.py", "reference_build": "oracle/Makefile",
                "compiler": flang, "fixtures": {}}
    jobs = [(n, lambda n=n, s=s, k=k: run_fixture(n, s, k)) for n, s, k in scenario_list()]
    sl = S.custom(laycord=0, laygov=1, nx=40, ny=36, **BEACH)
    cl = S.custom(laycord=1, laygov=1, nx=40, ny=36, **BEACH)
    ll = S.custom(laycord=0, laygov=0, nx=40, ny=36, cx=0.45)
    jobs += [
        ("local_mass_sph_nl", lambda: step_local("m", sl, "MASS", 1)),
        ("local_open_sph", lambda: step_local("o", ll, "OPEN", 2)),
        ("local_moment_sph_nl", lambda: step_local("x", sl, "MOMENT", 3)),
        ("local_moment_cart_nl", lambda: step_local("y", cl, "MOMENT", 4)),
        ("local_moment_sph_lin", lambda: step_local("z", ll, "MOMENT", 5)),
        ("local_moment_sph_brk", lambda: step_local_breaking(surf_zone(0), 6, 120)),
        ("local_moment_cart_brk", lambda: step_local_breaking(surf_zone(1), 7, 120)),
        ("local_moment_cart_brk_nofric",
         lambda: step_local_breaking(surf_zone(1, fric_coef=0.0), 8, 120)),
    ]
    for name, fn in jobs:
        arrays = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        manifest["fixtures"][name] = {"sha256": digest(arrays), "bytes": os.path.getsize(path)}
        print(name, os.path.getsize(path))
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
