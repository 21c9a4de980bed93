"""Seeded synthetic scenarios: the layer layouts of the benchmark configurations.

A scenario is a list of LayerSpec (root first, children in ALL_GRID order)
plus a step count.  Planes are numpy float32 arrays of shape (ny, nx) in C
order, which is the Fortran (NX, NY) column-major layout of the reference
(I fastest).  Depths and sources come from the library's generator
(csrc/synth.cpp), a pure function of (seed, point).

Workloads (seed = 20261202 + 1000*k + case, k = 0, 1, 2 in this order):
  linear_layer     spherical linear, 1024x768 at 2', 8 seamounts + shelf ramp,
           elliptical uplift 1 m (60x25 km), CFL 0.5, 2000 steps
  nonlinear_layer  spherical nonlinear, 4096x3072 at 0.5', + 3 embayments with
           1:50 beaches, dipole +-2 m (sigma 40 km, 80 km apart), Manning
           0.013, CFL 0.35, 5000 steps
  two_level_nest   spherical linear parent 2048x1536 at 2' with a Cartesian
           nonlinear child, REL_SIZE 4, REL_TIME 4, footprint 450x450
           parent cells over the shelf ramp, source 300 km offshore of the
           child, 3000 parent steps
All take size overrides so tests can run the same generator small.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import arr, fp, ip, lib

R_EARTH = 6378000.0
RAD_MIN = 0.000290888208665721
SEED0 = 20261202


class World(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_ulonglong),
                ("lon0", ctypes.c_double), ("lat0", ctypes.c_double),
                ("lon1", ctypes.c_double), ("lat1", ctypes.c_double),
                ("shelf_frac", ctypes.c_double), ("land_height", ctypes.c_double),
                ("ocean_depth", ctypes.c_double),
                ("n_seamounts", ctypes.c_int), ("n_embayments", ctypes.c_int),
                ("sm_lon", ctypes.c_double * 16), ("sm_lat", ctypes.c_double * 16),
                ("sm_height", ctypes.c_double * 16), ("sm_sigma_km", ctypes.c_double * 16),
                ("bay_lat", ctypes.c_double * 8), ("bay_halfwidth_km", ctypes.c_double * 8),
                ("bay_depth_km", ctypes.c_double * 8),
                ("beach_slope", ctypes.c_double), ("bay_floor", ctypes.c_double),
                ("coast_frac", ctypes.c_double)]


class Source(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("peak", ctypes.c_double),
                ("sig_a_km", ctypes.c_double), ("sig_b_km", ctypes.c_double),
                ("sep_km", ctypes.c_double), ("clon", ctypes.c_double),
                ("clat", ctypes.c_double), ("strike", ctypes.c_double)]


@dataclass
class LayerSpec:
    id: int
    parent: int
    level: int
    laycord: int          # 0 spherical, 1 Cartesian
    laygov: int           # 0 linear, 1 nonlinear
    nx: int
    ny: int
    dx: float             # arc minutes (spherical) or metres (Cartesian)
    dy: float
    dt: float
    h: np.ndarray         # (ny, nx) float32 still-water depth, < 0 on land
    z0: np.ndarray        # (ny, nx) float32 initial elevation
    y: np.ndarray = None  # (ny,) latitude of each row, degrees (spherical)
    dely: np.ndarray = None  # (ny,) row spacing, arc minutes (spherical)
    fric_switch: int = 1  # 0 Manning with fric_coef, 1 none
    fric_coef: float = 0.0
    rel_size: int = 1
    rel_time: int = 1
    corners: tuple = (0, 0, 0, 0)   # IS, IE, JS, JE in the parent (1-based)
    # SC_OPTION = 1 (Cartesian child of the spherical layer 1): POS (nx*ny*2)
    # and CXY (nx*ny*4) of SUBGRID_MATCHING, column- and component-major;
    # corners are then the POS extent
    sc_pos: np.ndarray = None
    sc_cxy: np.ndarray = None

    @property
    def cells(self):
        return self.nx * self.ny


@dataclass
class Scenario:
    name: str
    layers: list
    steps: int
    meta: dict = field(default_factory=dict)

    def cell_updates_per_step(self):
        """Cell updates of one root step: each layer times its sub-steps."""
        sub = {1: 1}
        total = 0
        for L in self.layers:
            s = 1 if L.parent == -1 else sub[L.parent] * L.rel_time
            sub[L.id] = s
            total += L.cells * s
        return total


def make_world(seed, lon0, lat0, lon1, lat1, n_seamounts=8, n_embayments=0,
               beach_slope=None, bay_floor=None, coast_frac=None):
    w = World()
    lib().gcomcot_synth_world(ctypes.c_ulonglong(seed), ctypes.c_double(lon0),
                              ctypes.c_double(lat0), ctypes.c_double(lon1),
                              ctypes.c_double(lat1), ctypes.c_int(n_seamounts),
                              ctypes.c_int(n_embayments), ctypes.byref(w))
    if beach_slope is not None:
        w.beach_slope = beach_slope
    if bay_floor is not None:
        w.bay_floor = bay_floor
    if coast_frac is not None:
        w.coast_frac = coast_frac
    return w


def depth(world, lon, lat):
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    h = np.empty(lon.shape, dtype=np.float32)
    lib().gcomcot_synth_depth(ctypes.byref(world), ctypes.c_long(lon.size),
                              lon.ctypes.data_as(ctypes.c_void_p),
                              lat.ctypes.data_as(ctypes.c_void_p),
                              h.ctypes.data_as(ctypes.c_void_p))
    return h


def source(world, kind, peak, sig_a_km, sig_b_km, sep_km, cx_frac, cy_frac, strike):
    s = Source()
    lib().gcomcot_synth_source(ctypes.byref(world), ctypes.c_int(kind), ctypes.c_double(peak),
                               ctypes.c_double(sig_a_km), ctypes.c_double(sig_b_km),
                               ctypes.c_double(sep_km), ctypes.c_double(cx_frac),
                               ctypes.c_double(cy_frac), ctypes.c_double(strike),
                               ctypes.byref(s))
    return s


def source_draw(world, stream, kind, peak, sig_a_km, sig_b_km, sep_km=0.0):
    s = Source()
    lib().gcomcot_synth_source_draw(ctypes.byref(world), ctypes.c_ulonglong(stream),
                                    ctypes.c_int(kind), ctypes.c_double(peak),
                                    ctypes.c_double(sig_a_km), ctypes.c_double(sig_b_km),
                                    ctypes.c_double(sep_km), ctypes.byref(s))
    return s


def eta(src, lon, lat, h):
    lon = np.ascontiguousarray(lon, dtype=np.float64)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float32)
    z = np.empty(lon.shape, dtype=np.float32)
    lib().gcomcot_synth_eta(ctypes.byref(src), ctypes.c_long(lon.size),
                            lon.ctypes.data_as(ctypes.c_void_p),
                            lat.ctypes.data_as(ctypes.c_void_p),
                            h.ctypes.data_as(ctypes.c_void_p),
                            z.ctypes.data_as(ctypes.c_void_p))
    return z


def cr_dt(laycord, laygov, h, dx, dy, lat_first, lat_last, dt_in=1.0e6):
    """CR_CHECK time step of a first-level layer (library restatement)."""
    ny, nx = h.shape
    ha, hp = arr(h)
    return float(lib().gcomcot_cr_dt_(ip(laycord), ip(laygov), ip(nx), ip(ny), hp, fp(dx), fp(dy),
                                      fp(lat_first), fp(lat_last), fp(dt_in)))


def para_rows(laygov, dt, dx, y, dely):
    """Per-row PARA table (10 x ny float32) of a spherical layer."""
    ny = len(y)
    ya, yp = arr(y)
    da, dp = arr(dely)
    coef = np.zeros(10 * ny, dtype=np.float32)
    lib().gcomcot_para_rows_(ip(laygov), ip(ny), fp(dt), fp(dx), yp, dp,
                             coef.ctypes.data_as(ctypes.c_void_p))
    return coef


def _sph_root(world, nx, ny, dx, laygov, src, fric_switch, fric_coef, lon0, lat0, dt=None):
    lon = (lon0 + np.arange(nx, dtype=np.float64) * dx / 60.0)
    lat = (lat0 + np.arange(ny, dtype=np.float64) * dx / 60.0)
    LON, LAT = np.meshgrid(lon, lat)
    h = depth(world, LON, LAT)
    z0 = eta(src, LON, LAT, h)
    y = lat.astype(np.float32)
    dely = np.full(ny, dx, dtype=np.float32)
    if dt is None:
        dt = cr_dt(0, laygov, h, dx, dx, float(y[0]), float(y[-1]))
    return LayerSpec(id=1, parent=-1, level=1, laycord=0, laygov=laygov, nx=nx, ny=ny,
                     dx=float(np.float32(dx)), dy=float(np.float32(dx)), dt=dt, h=h, z0=z0,
                     y=y, dely=dely, fric_switch=fric_switch, fric_coef=fric_coef)


def linear_layer(nx=1024, ny=768, dx=2.0, case=0, steps=2000, laygov=0):
    """Spherical linear single layer, 1024x768 at 2 arc-min by default."""
    lon0, lat0 = 140.0, 10.0
    lon1 = lon0 + (nx - 1) * dx / 60.0
    lat1 = lat0 + (ny - 1) * dx / 60.0
    seed = SEED0 + case
    w = make_world(seed, lon0, lat0, lon1, lat1, 8, 0)
    src = source_draw(w, 0, 0, 1.0, 60.0, 25.0)
    L = _sph_root(w, nx, ny, dx, laygov, src, 1, 0.0, lon0, lat0)
    return Scenario(f"linear-{nx}x{ny}", [L], steps, {"seed": seed})


def nonlinear_layer(nx=4096, ny=3072, dx=0.5, case=0, steps=5000, embayments=3):
    """Spherical nonlinear single layer with Manning friction, 4096x3072 at 0.5 arc-min by
    default."""
    lon0, lat0 = 140.0, 10.0
    lon1 = lon0 + (nx - 1) * dx / 60.0
    lat1 = lat0 + (ny - 1) * dx / 60.0
    seed = SEED0 + 1000 + case
    w = make_world(seed, lon0, lat0, lon1, lat1, 8, embayments)
    src = source(w, 1, 2.0, 40.0, 40.0, 80.0, 0.30, 0.5, 0.35)
    L = _sph_root(w, nx, ny, dx, 1, src, 0, 0.013, lon0, lat0)
    return Scenario(f"nonlinear-{nx}x{ny}", [L], steps, {"seed": seed})


def nonlinear_layer_sources(n, nx=4096, ny=3072, dx=0.5, case=0, steps=(1000, 4000),
                            embayments=3, seed=0):
    """n nonlinear layers of nonlinear_layer's world (one bathymetry, built
    once) with fresh dipole sources: peak 0.5-3 m, sigma 20-80 km, separation
    1-3 sigma, random position and strike; root step counts uniform in
    `steps`.  The layers share the depth array (read only)."""
    lon0, lat0 = 140.0, 10.0
    lon1 = lon0 + (nx - 1) * dx / 60.0
    lat1 = lat0 + (ny - 1) * dx / 60.0
    wseed = SEED0 + 1000 + case
    w = make_world(wseed, lon0, lat0, lon1, lat1, 8, embayments)
    src0 = source(w, 1, 2.0, 40.0, 40.0, 80.0, 0.30, 0.5, 0.35)
    base = _sph_root(w, nx, ny, dx, 1, src0, 0, 0.013, lon0, lat0)
    lon = lon0 + np.arange(nx, dtype=np.float64) * dx / 60.0
    lat = lat0 + np.arange(ny, dtype=np.float64) * dx / 60.0
    LON, LAT = np.meshgrid(lon, lat)
    rng = np.random.default_rng(SEED0 + 5000 + 97 * case + seed)
    out = []
    for k in range(n):
        sig = float(rng.uniform(20.0, 80.0))
        src = source(w, 1, float(rng.uniform(0.5, 3.0)), sig, sig, sig * float(rng.uniform(1.0, 3.0)),
                     float(rng.uniform(0.25, 0.8)), float(rng.uniform(0.2, 0.8)),
                     float(rng.uniform(0.0, math.pi)))
        L = LayerSpec(**{**base.__dict__, "z0": eta(src, LON, LAT, base.h)})
        out.append(Scenario(f"nonlinear-{nx}x{ny}-src{seed}.{k}", [L],
                            int(rng.integers(steps[0], steps[1] + 1)),
                            {"seed": wseed, "source": k}))
    return out


def two_level_nest(pnx=2048, pny=1536, dx=2.0, footprint=450, rel_size=4, rel_time=4, case=0,
            steps=3000, child_laygov=1, corner_i=30):
    """Spherical linear parent + Cartesian nonlinear child (2048x1536 + 1801x1801 by
    default)."""
    lon0, lat0 = 140.0, 10.0
    lon1 = lon0 + (pnx - 1) * dx / 60.0
    lat1 = lat0 + (pny - 1) * dx / 60.0
    seed = SEED0 + 2000 + case
    w = make_world(seed, lon0, lat0, lon1, lat1, 8, 0)
    IS, IE = corner_i, corner_i + footprint - 1
    JS = (pny - footprint) // 2 + 1
    JE = JS + footprint - 1
    dlon = dx / 60.0
    # child cell centres: ghost cell 1 west/south of the footprint
    cnx = footprint * rel_size + 1
    cny = cnx
    west = lon0 + (IS - 1) * dlon - 0.5 * dlon
    south = lat0 + (JS - 1) * dlon - 0.5 * dlon
    clon = west + (np.arange(1, cnx + 1) - 1.5) * dlon / rel_size
    clat = south + (np.arange(1, cny + 1) - 1.5) * dlon / rel_size
    mid_lat = 0.5 * (clat[0] + clat[-1])
    # source 300 km east of the child's east edge, at its middle latitude
    east = lon0 + (IE - 1) * dlon + 0.5 * dlon
    src_lon = east + 300.0 / (111.32 * math.cos(math.radians(mid_lat)))
    src = source(w, 0, 1.0, 60.0, 25.0, 0.0, (src_lon - lon0) / (lon1 - lon0),
                 (mid_lat - lat0) / (lat1 - lat0), 0.6)
    P = _sph_root(w, pnx, pny, dx, 0, src, 1, 0.0, lon0, lat0)
    CLON, CLAT = np.meshgrid(clon, clat)
    ch = depth(w, CLON, CLAT)
    cz = eta(src, CLON, CLAT, ch)
    cdx = float(np.float32(R_EARTH * math.cos(math.radians(mid_lat)) * dx * RAD_MIN / rel_size))
    cdy = float(np.float32(R_EARTH * dx * RAD_MIN / rel_size))
    cdt = float(np.float32(np.float32(P.dt) / np.float32(rel_time)))
    C = LayerSpec(id=2, parent=1, level=2, laycord=1, laygov=child_laygov, nx=cnx, ny=cny,
                  dx=cdx, dy=cdy, dt=cdt, h=ch, z0=cz, fric_switch=0 if child_laygov else 1,
                  fric_coef=0.013 if child_laygov else 0.0, rel_size=rel_size,
                  rel_time=rel_time, corners=(IS, IE, JS, JE))
    return Scenario(f"nest-{pnx}x{pny}+{cnx}x{cny}", [P, C], steps, {"seed": seed})


def layer_coords(s):
    """Grid coordinates X(NX), Y(NY) of a layer (the LAYER%X / %Y that gauges
    are located on), float32: spherical layers use degrees east of their
    first column and their row latitudes; Cartesian layers metres from
    their first cell."""
    if s.laycord == 0:
        x = (np.arange(s.nx, dtype=np.float64) * s.dx / 60.0).astype(np.float32)
        y = np.asarray(s.y, np.float32)
    else:
        x = (np.arange(s.nx, dtype=np.float64) * s.dx).astype(np.float32)
        y = (np.arange(s.ny, dtype=np.float64) * s.dy).astype(np.float32)
    return x, y


def child_dt(laycord, laygov, h, dx, dy, lat_first, lat_last, dt_parent):
    """CR_CHECK for a nested layer (initialization.f90:1206-1247): (REL_TIME, DT)."""
    ha, hp = arr(h)
    rt = ctypes.c_int(0)
    dt = ctypes.c_float(0)
    lib().gcomcot_child_dt_(ip(laycord), ip(laygov), ip(h.shape[1]), ip(h.shape[0]), hp,
                            fp(dx), fp(dy), fp(lat_first), fp(lat_last), fp(dt_parent),
                            ctypes.byref(rt), ctypes.byref(dt))
    return rt.value, float(dt.value)


def _cart_child(world, src, parent, dx, lon0, lat0, IS, JS, footprint_i, footprint_j,
                rel_size, laygov=1, rel_time=None):
    """Cartesian child over parent cells IS.., JS.. (ghost row/column first)."""
    dlon = dx / 60.0
    IE, JE = IS + footprint_i - 1, JS + footprint_j - 1
    cnx = footprint_i * rel_size + 1
    cny = footprint_j * rel_size + 1
    west = lon0 + (IS - 1) * dlon - 0.5 * dlon
    south = lat0 + (JS - 1) * dlon - 0.5 * dlon
    clon = west + (np.arange(1, cnx + 1) - 1.5) * dlon / rel_size
    clat = south + (np.arange(1, cny + 1) - 1.5) * dlon / rel_size
    mid_lat = 0.5 * (clat[0] + clat[-1])
    CLON, CLAT = np.meshgrid(clon, clat)
    ch = depth(world, CLON, CLAT)
    cz = eta(src, CLON, CLAT, ch)
    cdx = float(np.float32(R_EARTH * math.cos(math.radians(mid_lat)) * dx * RAD_MIN / rel_size))
    cdy = float(np.float32(R_EARTH * dx * RAD_MIN / rel_size))
    if rel_time is None:
        rel_time, cdt = child_dt(1, laygov, ch, cdx, cdy, 0.0, 0.0, parent.dt)
    else:
        cdt = float(np.float32(np.float32(parent.dt) / np.float32(rel_time)))
    return LayerSpec(id=2, parent=1, level=2, laycord=1, laygov=laygov, nx=cnx, ny=cny,
                     dx=cdx, dy=cdy, dt=cdt, h=ch, z0=cz, fric_switch=0 if laygov else 1,
                     fric_coef=0.013 if laygov else 0.0, rel_size=rel_size,
                     rel_time=rel_time, corners=(IS, IE, JS, JE))


def scenario_mix(n=64, case=0, min_cells=256 * 256, max_cells=2048 * 1536, nest_share=0.25,
            steps=(1000, 4000), child_side=(400, 1200)):
    """Mix of independent scenarios for the batch engine and the serving loop.

    Parent cell counts log-uniform in [min_cells, max_cells], aspect 4:3,
    spacing 1-4 arc minutes, spherical linear; a share `nest_share` of the
    cases carries one Cartesian nonlinear child (REL_SIZE 3-5, child side
    `child_side` cells, REL_TIME from CR_CHECK) over the shelf ramp;
    elliptical sources 0.5-3 m, sigma 20-80 km, random position and
    strike; root step counts uniform in `steps`.  Returns a list of
    Scenario, each with its own `steps`."""
    rng = np.random.default_rng(SEED0 + 3000 + case)
    out = []
    for k in range(n):
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        ny = max(24, int(round(math.sqrt(cells * 3.0 / 4.0))))
        nx = max(32, int(round(ny * 4.0 / 3.0)))
        # spacing 1-4 arc minutes, coarsest first dropped until the domain
        # spans at most 50 degrees of latitude (and stays within +-65)
        fits = [d for d in (1.0, 2.0, 3.0, 4.0) if (ny - 1) * d / 60.0 <= 50.0]
        dx = float(rng.choice(fits))
        nested = (k % round(1.0 / nest_share) == 0) if nest_share > 0 else False
        span = (ny - 1) * dx / 60.0
        lon0, lat0 = 140.0, float(rng.uniform(-10.0, 65.0 - span))
        lon1 = lon0 + (nx - 1) * dx / 60.0
        lat1 = lat0 + (ny - 1) * dx / 60.0
        seed = SEED0 + 3000 + 100 * case + k
        w = make_world(seed, lon0, lat0, lon1, lat1, 8, 0)
        sa = float(rng.uniform(20.0, 80.0))
        sb = float(rng.uniform(20.0, sa))
        src = source(w, 0, float(rng.uniform(0.5, 3.0)), sa, sb, 0.0,
                     float(rng.uniform(0.35, 0.8)), float(rng.uniform(0.2, 0.8)),
                     float(rng.uniform(0.0, math.pi)))
        P = _sph_root(w, nx, ny, dx, 0, src, 1, 0.0, lon0, lat0)
        layers = [P]
        if nested:
            R = int(rng.choice([3, 4, 5]))
            side = int(rng.uniform(child_side[0], child_side[1]))
            fp_ = max(4, min(side // R, ny - 12, nx // 2))
            IS = max(4, int(0.02 * nx))
            JS = max(4, (ny - fp_) // 2 + 1)
            layers.append(_cart_child(w, src, P, dx, lon0, lat0, IS, JS, fp_, fp_, R))
        st = int(rng.integers(steps[0], steps[1] + 1))
        name = f"mix-{k}-{nx}x{ny}" + (f"+{layers[1].nx}x{layers[1].ny}" if nested else "")
        out.append(Scenario(name, layers, st, {"seed": seed}))
    return out


def custom(nx=64, ny=56, dx=2.0, laycord=0, laygov=0, fric_coef=0.013, child=None, case=0,
           steps=100, embayments=1, peak=1.5, cx=0.45, cy=0.5, strike=0.4, src_kind=0,
           sig=40.0, beach_slope=None, bay_floor=None, coast_frac=None):
    """Small scenario of any layer kind, for coverage tests.

    laycord 1 roots take dx in arc minutes too: the grid is laid on the
    same lon/lat world and dx is converted to metres at its mid-latitude.
    child: None or dict(laycord, laygov, footprint=(w, h), rel_size,
    rel_time, corner=(IS, JS)).
    """
    lon0, lat0 = 140.0, 10.0
    lon1 = lon0 + (nx - 1) * dx / 60.0
    lat1 = lat0 + (ny - 1) * dx / 60.0
    seed = SEED0 + 9000 + case
    w = make_world(seed, lon0, lat0, lon1, lat1, 8, embayments, beach_slope, bay_floor,
                   coast_frac)
    src = source(w, src_kind, peak, sig, 0.6 * sig, 1.5 * sig, cx, cy, strike)
    fs = 0 if fric_coef > 0 else 1
    if laycord == 0:
        P = _sph_root(w, nx, ny, dx, laygov, src, fs, fric_coef, lon0, lat0)
    else:
        lon = lon0 + np.arange(nx, dtype=np.float64) * dx / 60.0
        lat = lat0 + np.arange(ny, dtype=np.float64) * dx / 60.0
        LON, LAT = np.meshgrid(lon, lat)
        h = depth(w, LON, LAT)
        z0 = eta(src, LON, LAT, h)
        mid = 0.5 * (lat[0] + lat[-1])
        dxm = float(np.float32(R_EARTH * math.cos(math.radians(mid)) * dx * RAD_MIN))
        dym = float(np.float32(R_EARTH * dx * RAD_MIN))
        dt = cr_dt(1, laygov, h, dxm, dym, 0.0, 0.0)
        P = LayerSpec(id=1, parent=-1, level=1, laycord=1, laygov=laygov, nx=nx, ny=ny, dx=dxm,
                      dy=dym, dt=dt, h=h, z0=z0, fric_switch=fs, fric_coef=fric_coef)
    layers = [P]
    if child:
        R, RT = child["rel_size"], child["rel_time"]
        fw, fh = child["footprint"]
        IS, JS = child["corner"]
        IE, JE = IS + fw - 1, JS + fh - 1
        dlon = dx / 60.0
        cnx, cny = fw * R + 1, fh * R + 1
        west = lon0 + (IS - 1) * dlon - 0.5 * dlon
        south = lat0 + (JS - 1) * dlon - 0.5 * dlon
        clon = west + (np.arange(1, cnx + 1) - 1.5) * dlon / R
        clat = south + (np.arange(1, cny + 1) - 1.5) * dlon / R
        CLON, CLAT = np.meshgrid(clon, clat)
        ch = depth(w, CLON, CLAT)
        cz = eta(src, CLON, CLAT, ch)
        cdt = float(np.float32(np.float32(P.dt) / np.float32(RT)))
        clc, clg = child["laycord"], child["laygov"]
        cfs = 0 if (fric_coef > 0 and clg == 1) else 1
        if clc == 0:
            C = LayerSpec(id=2, parent=1, level=2, laycord=0, laygov=clg, nx=cnx, ny=cny,
                          dx=float(np.float32(dx / R)), dy=float(np.float32(dx / R)), dt=cdt,
                          h=ch, z0=cz, y=clat.astype(np.float32),
                          dely=np.full(cny, np.float32(dx / R), np.float32), fric_switch=cfs,
                          fric_coef=fric_coef if cfs == 0 else 0.0, rel_size=R, rel_time=RT,
                          corners=(IS, IE, JS, JE))
        else:
            mid = 0.5 * (clat[0] + clat[-1])
            cdx = float(np.float32(R_EARTH * math.cos(math.radians(mid)) * dx * RAD_MIN / R))
            cdy = float(np.float32(R_EARTH * dx * RAD_MIN / R))
            C = LayerSpec(id=2, parent=1, level=2, laycord=1, laygov=clg, nx=cnx, ny=cny, dx=cdx,
                          dy=cdy, dt=cdt, h=ch, z0=cz, fric_switch=cfs,
                          fric_coef=fric_coef if cfs == 0 else 0.0, rel_size=R, rel_time=RT,
                          corners=(IS, IE, JS, JE))
        layers.append(C)
    return Scenario(f"custom-{laycord}{laygov}-{nx}x{ny}", layers, steps, {"seed": seed})


def bc_wall(h):
    """BC_WALL (boundaries.f90:369-401) through the library: a copy of the
    depth h (ny, nx) with its two outermost rows and columns turned to land."""
    import ctypes
    from ._lib import lib
    out = np.ascontiguousarray(h, np.float32).copy()
    ny, nx = out.shape
    lib().gcomcot_bc_wall_(out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_int(nx)),
                           ctypes.byref(ctypes.c_int(ny)))
    return out
