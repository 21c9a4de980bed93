"""CPU checks of the pieces around the step: ABI exports, the scalar
pow/exp, the scenario generator and invariants the mathematics fixes."""
import ctypes
import os
import re

import numpy as np
import pytest

from gcomcot import scenarios as S
from helpers import bits
from oracle.bind import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gcomcot_hip.h")

# the 19 symbols the reference Fortran objects leave undefined for the GPU
# half (comcot.f90, output.f90, all_grids.f90)
REFERENCE_SYMBOLS = """all_grid_launch_ cuda_getmn1_ cuda_getmn_ cuda_getz1_ cuda_getz_
cuda_getzmax_ cuda_setmn1_ cuda_setmn_ cuda_setz1_ cuda_setz_ cuda_update_layer_ gcomcot_get_z_
gcomcot_init_gpu_ gcomcot_init_layer_ gcomcot_show_grids_ mass_launch_ maxamp_launch_
momt_launch_ openbd_launch_""".split()


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:void|int|float|long long)\s+(\w+)\s*\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    from gcomcot import lib
    L = lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/gcomcot_hip.h but not exported"


def test_reference_boundary_is_complete():
    names = set(header_functions())
    missing = [s for s in REFERENCE_SYMBOLS if s not in names]
    assert not missing, missing


FLANG = "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_interface_module_compiles(tmp_path):
    import subprocess
    src = os.path.join(ROOT, "include", "gcomcot_hip.f90")
    subprocess.run([FLANG, "-c", src, "-o", str(tmp_path / "m.o"), "-module-dir",
                    str(tmp_path)], check=True)
    # every bind(C) name of the module is a symbol the library exports
    names = re.findall(r'bind\(C, name="(\w+)"\)', open(src).read())
    from gcomcot import lib
    L = lib()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.isdir("/root/reference")),
                    reason="needs flang and the reference sources")
def test_unmodified_reference_links_against_the_library():
    """The reference's own Fortran, main program included, compiled as is
    and linked against libgcomcot_hip.so: every GPU entry it calls must
    resolve to this library (the link of the reference Makefile:38)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin",
                    "REF=/root/reference"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin", "comcot_hip")
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], check=True,
                         capture_output=True, text=True).stdout
    undefined = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert set(REFERENCE_SYMBOLS) <= undefined  # taken from the shared library
    ldd = subprocess.run(["ldd", exe], check=True, capture_output=True, text=True).stdout
    assert "libgcomcot_hip.so" in ldd and "not found" not in ldd.split("libgcomcot_hip")[1][:80]


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.isdir("/root/reference")),
                    reason="needs flang and the reference sources")
def test_reference_debug_build_links_against_the_library():
    """The reference compiled with its differential-debug switch (DEBUG_CORE,
    debug_option.h) also calls edgeinterp_dbglaunch_ and jnz_dbglaunch_
    (all_grids.f90:417, 490, 650): that build must link too."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin-debug",
                    "REF=/root/reference"], check=True, capture_output=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin_debug", "comcot_hip_debug")
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], check=True,
                         capture_output=True, text=True).stdout
    undefined = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert {"edgeinterp_dbglaunch_", "jnz_dbglaunch_", "mass_launch_", "momt_launch_"} <= undefined


def _libm():
    return ctypes.CDLL("libm.so.6")


def test_det_pow_is_correctly_rounded(oracle_built):
    """powf_det against the correctly rounded x^(2.3333333f) (float64 pow of
    exact float inputs), over the depth range the momentum routines use."""
    L = Oracle.lib()
    L.orc_powf_det.restype = ctypes.c_float
    L.orc_powf_det.argtypes = [ctypes.c_float, ctypes.c_float]
    rng = np.random.default_rng(7)
    xs = np.exp(rng.uniform(np.log(1e-3), np.log(8000.0), 20000)).astype(np.float32)
    y = np.float32(2.3333333)
    det = np.array([L.orc_powf_det(float(x), float(y)) for x in xs], np.float32)
    cr = np.power(xs.astype(np.float64), np.float64(y)).astype(np.float32)
    assert np.array_equal(bits(det), bits(cr))
    # libm's powf agrees except on a small fraction (it is not correctly rounded)
    m = _libm()
    m.powf.restype = ctypes.c_float
    m.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    lm = np.array([m.powf(float(x), float(y)) for x in xs], np.float32)
    assert np.mean(lm != det) < 2e-3


def test_manning_pow_equals_det_pow_on_its_domain(oracle_built):
    """powf_manning (the branch-free form the kernels call) returns the bits
    of powf_det for every depth the momentum routines can pass (>= 1e-5),
    up to the largest float, and for +inf and NaN."""
    L = Oracle.lib()
    for f in (L.orc_powf_det, L.orc_powf_manning):
        f.restype = ctypes.c_float
        f.argtypes = [ctypes.c_float, ctypes.c_float]
    rng = np.random.default_rng(11)
    xs = np.exp(rng.uniform(np.log(1e-5), np.log(3.4e38), 40000)).astype(np.float32)
    edges = np.array([1e-5, 1e-3, 1.0, 3.0e38, 3.0000002e38, 3.4028235e38, np.inf, np.nan],
                     np.float32)
    y = float(np.float32(2.3333333))
    for x in np.concatenate([xs, edges]):
        a = L.orc_powf_det(float(x), y)
        b = L.orc_powf_manning(float(x), y)
        assert bits(np.float32(a)) == bits(np.float32(b)) or (np.isnan(a) and np.isnan(b)), x


def test_det_exp_matches_libm_on_edges(oracle_built):
    L = Oracle.lib()
    L.orc_expf_det.restype = ctypes.c_float
    L.orc_expf_det.argtypes = [ctypes.c_float]
    m = _libm()
    m.expf.restype = ctypes.c_float
    m.expf.argtypes = [ctypes.c_float]
    for x in (-np.inf, -1e30, -200.0, -104.0, -103.0, -90.0, -20.0, -1.5, 0.0, 3.0, 88.7, 89.0,
              np.inf):
        a, b = L.orc_expf_det(x), m.expf(x)
        assert a == b, (x, a, b)
    assert np.isnan(L.orc_expf_det(float("nan")))


def test_generator_is_deterministic():
    a = S.linear_layer(nx=64, ny=48)
    b = S.linear_layer(nx=64, ny=48)
    L1, L2 = a.layers[0], b.layers[0]
    assert np.array_equal(bits(L1.h), bits(L2.h)) and np.array_equal(bits(L1.z0), bits(L2.z0))
    assert L1.dt == L2.dt
    c = S.linear_layer(nx=64, ny=48, case=1)
    assert not np.array_equal(L1.z0, c.layers[0].z0)


def test_generator_world_shape():
    sc = S.nonlinear_layer(nx=512, ny=384, dx=4.0)
    L = sc.layers[0]
    h = L.h
    assert (h[:, 0] < 0).all(), "west edge is dry land"
    assert (h[:, -1] > 1000).all(), "east edge is deep ocean"
    assert 0.01 < np.mean(h <= 0) < 0.2
    # dipole of two 2.0 m Gaussian lobes (sigma 40 km) 80 km apart: the
    # overlap lowers each extremum to about 2*(1-exp(-2)) ~ 1.73..1.8 m
    assert 1.6 < float(L.z0.max()) <= 2.0 and -2.0 <= float(L.z0.min()) < -1.6


@pytest.mark.parametrize("laycord,laygov", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_lake_at_rest_stays_at_rest(laycord, laygov, oracle_built):
    """Z = M = N = 0 over any bathymetry is a fixed point of every form
    (dry cells of the nonlinear forms sit at Z = -H)."""
    sc = S.custom(laycord=laycord, laygov=laygov, nx=40, ny=36, peak=0.0, **{
        "embayments": 2, "beach_slope": 3000.0, "bay_floor": 6.0, "coast_frac": 0.3})
    L = sc.layers[0]
    o = Oracle(sc, 32, 0)
    o.step(30)
    st = o.state(1)
    assert not st["m1"].any() and not st["n1"].any()
    if laygov == 0:
        assert not st["z1"].any()
    else:
        wet = L.h > 1e-5
        assert not st["z1"][wet].any()
        dry = (L.h < -1e-5)
        core = np.zeros_like(dry)
        core[1:-1, 1:-1] = True   # the open-boundary ring is not a continuity cell
        assert np.array_equal(st["z1"][dry & core], -L.h[dry & core])


def test_jnz_of_uniform_child_is_that_constant(oracle_built):
    sc = S.custom(nx=40, ny=36, laygov=0, peak=0.0, child=dict(
        laycord=1, laygov=0, footprint=(8, 9), rel_size=3, rel_time=3, corner=(20, 10)))
    o = Oracle(sc, 32, 0)
    C = sc.layers[1]
    st = o.state(2)
    st["z2"] = np.full_like(st["z2"], 0.375)
    st["z1"] = st["z2"].copy()
    o.set_state(2, st)
    o.call("JNZ", 2)
    p = o.state(1)["z2"]
    IS, IE, JS, JE = C.corners
    wet_child = (C.h[1:, 1:] + 0.375 > 1e-5)
    blk = p[JS - 1:JE, IS - 1:IE]
    # every parent cell whose child block is more than half wet gets exactly 0.375
    frac = wet_child.reshape(JE - JS + 1, 3, IE - IS + 1, 3).mean(axis=(1, 3))
    assert np.all(blk[frac > 0.5] == np.float32(0.375))
    assert np.all(blk[frac <= 0.5] == 0)


def test_jnq_reproduces_an_affine_flux(oracle_built):
    """EDGEINTERP reproduces a parent flux that is affine along the interface."""
    sc = S.custom(nx=40, ny=36, laygov=0, peak=0.0, coast_frac=0.0, child=dict(
        laycord=1, laygov=0, footprint=(8, 9), rel_size=4, rel_time=4, corner=(20, 10)))
    o = Oracle(sc, 32, 0)
    P, C = sc.layers
    st = o.state(1)
    jj = np.arange(1, P.ny + 1, dtype=np.float32)
    st["m1"] = np.repeat((0.25 * jj + 1.0)[:, None], P.nx, 1).astype(np.float32)
    o.set_state(1, st)
    o.call("JNQ", 2)
    cm = o.state(2)["m1"]
    IS, IE, JS, JE = C.corners
    # consecutive child faces step by the parent slope / R, and the first
    # face of parent row JS (child JJ = 2) carries the value at JS - 1/2 + 1/(2R)
    d = np.diff(cm[1:, 0].astype(np.float64))
    assert np.allclose(d, 0.25 / 4.0, atol=1e-6)
    assert abs(float(cm[1, 0]) - (0.25 * (JS - 0.5 + 1.0 / 8.0) + 1.0)) < 1e-5
