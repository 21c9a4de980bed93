"""Python bindings of the two CPU checkers (test infrastructure only).

  Oracle     oracle/_build/liboracle.so   C++ restatement (oracle.cpp),
             FP32 or FP64, libm or device pow/exp
  Reference  oracle/_ref/libcomcot_ref.so the reference's own Fortran,
             flang-built from its sources by oracle/Makefile

Both take a gcomcot Scenario (root first, children in id order 2, 3, ...)
and expose the same calls: step, call (one routine), state, set_state.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module.
"""
import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcomcot_ref.so")
REFERENCE_DIR = "/root/reference"

OP = dict(MASS=1, OPEN=2, MOMENT=3, JNQ=4, NEWQ=5, JNZ=6, UPDATE=7, ALL_GRID=8, CHANGE=9, MAX=10)

# Branches of CONMOME(_S) / WAVE_BREAKING counted by the oracle (enum Branch
# of oracle.cpp, same order)
BRANCHES = ("noflux", "dd_wet", "dd_edge_a", "dd_edge_b", "df_clamp", "fric_on", "fric_off",
            "adv", "up_neg_deg", "up_neg_full", "up_pos_deg", "up_pos_full", "cross_neg_dz",
            "cross_neg_dp", "cross_neg_full", "cross_pos_dz", "cross_pos_dp", "cross_pos_full",
            "brk_none", "brk_finite", "brk_over", "brk_dry", "vlim_hi", "vlim_lo", "eps_zero")


def build(ref=True):
    """Build liboracle.so, and the reference library when its sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle"], check=True)
    if ref and os.path.isdir(REFERENCE_DIR):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REFERENCE_DIR}"], check=True)


def big_stack(fn, *args, stack_gb=4):
    """Run fn on a thread with a large stack: the reference's momentum
    routines hold ~11 whole-plane automatic arrays on the stack."""
    out = {}

    def target():
        try:
            out["r"] = fn(*args)
        except BaseException as e:  # re-raised in the caller
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(stack_gb * 1024 ** 3 - 4096)
    try:
        th = threading.Thread(target=target)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _f(v):
    return ctypes.byref(ctypes.c_float(float(v)))


class Oracle:
    """The C++ restatement.  precision 32 or 64; pow_mode 0 libm, 1 device."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                build(ref=False)
            L = ctypes.CDLL(ORACLE_SO)
            L.orc_new.restype = ctypes.c_void_p
            L.orc_new.argtypes = [ctypes.c_int, ctypes.c_int]
            L.orc_free.argtypes = [ctypes.c_void_p]
            L.orc_add_layer.restype = ctypes.c_int
            L.orc_add_layer.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 6 + \
                [ctypes.c_double] * 3 + [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_double,
                                                                  ctypes.c_int, ctypes.c_int,
                                                                  ctypes.c_void_p]
            L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
            L.orc_call.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            L.orc_state_io.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + \
                [ctypes.c_void_p] * 4
            L.orc_setup_io.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5
            L.orc_flux_io.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 2
            L.orc_max_io.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
            L.orc_set_arrival_threshold.argtypes = [ctypes.c_void_p, ctypes.c_double]
            L.orc_set_sponge.argtypes = [ctypes.c_void_p] * 3
            L.orc_set_sc.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3
            L.orc_branch_counts.restype = ctypes.c_int
            L.orc_branch_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            L.orc_cr_dt.restype = ctypes.c_double
            L.orc_cr_dt.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] + [ctypes.c_double] * 5
            cls._lib = L
        return cls._lib

    def __init__(self, scenario, precision=32, pow_mode=0, arrival_threshold=0.01):
        L = self.lib()
        self.h = L.orc_new(precision, pow_mode)
        self.dtype = np.float32 if precision == 32 else np.float64
        self.specs = {}
        L.orc_set_arrival_threshold(self.h, arrival_threshold)
        for s in scenario.layers:
            self.specs[s.id] = s
            y = np.zeros(s.ny) if s.y is None else np.asarray(s.y, dtype=np.float64)
            dely = np.zeros(s.ny) if s.dely is None else np.asarray(s.dely, dtype=np.float64)
            h = np.ascontiguousarray(s.h, dtype=np.float64)
            z0 = np.ascontiguousarray(s.z0, dtype=np.float64)
            corners = np.array(s.corners, dtype=np.int32)
            rc = L.orc_add_layer(self.h, s.id, s.parent, s.laycord, s.laygov, s.nx, s.ny,
                                 s.dx, s.dy, s.dt, _p(y), _p(dely), _p(h), _p(z0),
                                 s.fric_switch, s.fric_coef, s.rel_size, s.rel_time, _p(corners))
            if rc != 0:
                raise ValueError(f"orc_add_layer({s.id}) = {rc}")
            if getattr(s, "sc_pos", None) is not None:
                self.set_sc(s.id, s.sc_pos, s.sc_cxy, s.corners)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().orc_free(self.h)
            self.h = None

    def step(self, n=1):
        self.lib().orc_step(self.h, n)

    def branch_counts(self):
        """Flux points per CONMOME(_S)/WAVE_BREAKING branch since the last
        call, {name: (x_count, y_count)}; the counters are cleared."""
        n = len(BRANCHES)
        out = np.zeros(2 * n, np.int64)
        got = self.lib().orc_branch_counts(self.h, _p(out))
        if got != n:
            raise RuntimeError(f"oracle has {got} branch counters, bind.py names {n}")
        return {k: (int(out[i]), int(out[n + i])) for i, k in enumerate(BRANCHES)}

    def set_sc(self, lid, pos, cxy, corners):
        """Couple child lid with SC_OPTION = 1: POS (nx*ny*2) and CXY (nx*ny*4)
        component-major as SUBGRID_MATCHING leaves them, and its corners."""
        p = np.ascontiguousarray(pos, np.int32).reshape(-1)
        c = np.ascontiguousarray(cxy, np.float64).reshape(-1)
        k = np.ascontiguousarray(corners, np.int32)
        if self.lib().orc_set_sc(self.h, lid, _p(p), _p(c), _p(k)) != 0:
            raise ValueError(f"orc_set_sc({lid})")

    def set_sponge(self, cx, cy):
        """SPONGE_LAYER after every momentum pass of layer 1, with COEFX per
        column (nx) and COEFY per row (ny)."""
        a = np.ascontiguousarray(cx, np.float64)
        b = np.ascontiguousarray(cy, np.float64)
        self.lib().orc_set_sponge(self.h, _p(a), _p(b))

    def call(self, op, lid=1, t=0):
        self.lib().orc_call(self.h, OP[op] if isinstance(op, str) else op, lid, t)

    def state(self, lid=1):
        s = self.specs[lid]
        bufs = [np.empty(2 * s.nx * s.ny) for _ in range(4)]
        self.lib().orc_state_io(self.h, lid, 0, *[_p(b) for b in bufs])
        out = {}
        for name, b in zip(("z", "m", "n", "dz"), bufs):
            b = b.reshape(2, s.ny, s.nx).astype(self.dtype)
            out[name + "1"], out[name + "2"] = b[0], b[1]
        return out

    def set_state(self, lid, st):
        s = self.specs[lid]
        bufs = []
        for name in ("z", "m", "n", "dz"):
            b = np.stack([np.asarray(st[name + "1"], np.float64), np.asarray(st[name + "2"], np.float64)])
            bufs.append(np.ascontiguousarray(b.reshape(-1)))
        self.lib().orc_state_io(self.h, lid, 1, *[_p(b) for b in bufs])

    def setup(self, lid=1):
        s = self.specs[lid]
        n = s.nx * s.ny
        hp, hq, rpl, r6, sc = (np.empty(n), np.empty(n), np.empty(9 * n), np.empty(s.ny),
                               np.empty(5))
        self.lib().orc_setup_io(self.h, lid, _p(hp), _p(hq), _p(rpl), _p(r6), _p(sc))
        return dict(hp=hp.reshape(s.ny, s.nx), hq=hq.reshape(s.ny, s.nx),
                    r=rpl.reshape(9, s.ny, s.nx), r6=r6, rx=sc[0], ry=sc[1], grx=sc[2],
                    gry=sc[3], dt=sc[4])

    def flux(self, lid=1):
        s = self.specs[lid]
        xf, yf = np.empty(2 * s.nx), np.empty(2 * s.ny)
        self.lib().orc_flux_io(self.h, lid, _p(xf), _p(yf))
        return xf.reshape(2, s.nx).astype(self.dtype), yf.reshape(2, s.ny).astype(self.dtype)

    def maxima(self, lid=1):
        s = self.specs[lid]
        n = s.nx * s.ny
        zw, zr, ar = np.empty(n), np.empty(n), np.empty(n, dtype=np.int32)
        self.lib().orc_max_io(self.h, lid, _p(zw), _p(zr), _p(ar))
        return {"zmax_wet": zw.reshape(s.ny, s.nx).astype(self.dtype),
                "zmax_raw": zr.reshape(s.ny, s.nx).astype(self.dtype),
                "arrival": ar.reshape(s.ny, s.nx)}


class Reference:
    """The flang-built reference Fortran (needs oracle/_ref built)."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(REF_SO):
                raise FileNotFoundError(REF_SO)
            cls._lib = ctypes.CDLL(REF_SO)
        return cls._lib

    def __init__(self, scenario):
        L = self.lib()
        self.specs = {}
        for k, s in enumerate(scenario.layers):
            if s.id != k + 1:
                raise ValueError("reference driver needs layer ids 1, 2, 3, ... in order")
            self.specs[s.id] = s
            y = np.zeros(s.ny, np.float32) if s.y is None else np.asarray(s.y, np.float32)
            dely = np.zeros(s.ny, np.float32) if s.dely is None else np.asarray(s.dely, np.float32)
            h = np.ascontiguousarray(s.h, dtype=np.float32)
            z0 = np.ascontiguousarray(s.z0, dtype=np.float32)
            self._keep = (y, dely, h, z0)
            if s.parent == -1:
                L.ref_init_root(_i(s.nx), _i(s.ny), _i(s.laycord), _i(s.laygov), _f(s.dx),
                                _f(s.dy), _f(s.dt), _p(y), _p(dely), _p(h), _p(z0),
                                _i(s.fric_switch), _f(s.fric_coef))
            else:
                corners = np.array(s.corners, dtype=np.int32)
                L.ref_init_child(_i(s.parent), _i(s.laycord), _i(s.laygov), _i(s.rel_size),
                                 _i(s.rel_time), _p(corners), _i(s.nx), _i(s.ny), _f(s.dx),
                                 _f(s.dy), _f(s.dt), _p(y), _p(dely), _p(h), _p(z0),
                                 _i(s.fric_switch), _f(s.fric_coef))

    def step(self, n=1):
        big_stack(self.lib().ref_step, _i(n))

    def call(self, op, lid=1, t=0):
        big_stack(self.lib().ref_call, _i(OP[op] if isinstance(op, str) else op), _i(lid), _i(t))

    def state(self, lid=1):
        s = self.specs[lid]
        bufs = [np.empty(2 * s.nx * s.ny, np.float32) for _ in range(4)]
        self.lib().ref_state_io(_i(lid), _i(0), *[_p(b) for b in bufs])
        out = {}
        for name, b in zip(("z", "m", "n", "dz"), bufs):
            b = b.reshape(2, s.ny, s.nx)
            out[name + "1"], out[name + "2"] = b[0].copy(), b[1].copy()
        return out

    def set_state(self, lid, st):
        bufs = []
        for name in ("z", "m", "n", "dz"):
            b = np.stack([np.asarray(st[name + "1"], np.float32), np.asarray(st[name + "2"], np.float32)])
            bufs.append(np.ascontiguousarray(b.reshape(-1)))
        self.lib().ref_state_io(_i(lid), _i(1), *[_p(b) for b in bufs])

    def setup(self, lid=1):
        s = self.specs[lid]
        n = s.nx * s.ny
        hp, hq, rpl = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(9 * n, np.float32)
        r6, sc = np.empty(s.ny, np.float32), np.empty(5, np.float32)
        self.lib().ref_setup_io(_i(lid), _p(hp), _p(hq), _p(rpl), _p(r6), _p(sc))
        return dict(hp=hp.reshape(s.ny, s.nx), hq=hq.reshape(s.ny, s.nx),
                    r=rpl.reshape(9, s.ny, s.nx), r6=r6, rx=sc[0], ry=sc[1], grx=sc[2],
                    gry=sc[3], dt=sc[4])

    def set_coords(self, lid, x, y):
        """LAYER%X(NX), %Y(NY) of a layer, the coordinates GET_TS searches."""
        xa = np.ascontiguousarray(x, np.float32)
        ya = np.ascontiguousarray(y, np.float32)
        self.lib().ref_set_coords(_i(lid), _p(xa), _p(ya))

    def get_ts(self, x0, y0, lid):
        """The reference's gauge sample GET_TS (output.f90:1035-1155) of the
        current elevation of layer lid at (x0, y0)."""
        dat = ctypes.c_float(0)
        self.lib().ref_get_ts(_f(x0), _f(y0), _i(lid), ctypes.byref(dat))
        return np.float32(dat.value)

    def flux(self, lid=1):
        s = self.specs[lid]
        xf, yf = np.empty(2 * s.nx, np.float32), np.empty(2 * s.ny, np.float32)
        self.lib().ref_flux_io(_i(lid), _p(xf), _p(yf))
        return xf.reshape(2, s.nx), yf.reshape(2, s.ny)

    def maxamp(self):
        """MAXAMP_LAUNCH after a root step: the recorder maximum of layer 1."""
        self.lib().ref_maxamp()

    def raw_max(self):
        s = self.specs[1]
        out = np.empty(s.nx * s.ny, np.float32)
        self.lib().cuda_getzmax_(_p(out), _i(1))
        return out.reshape(s.ny, s.nx)

    def data_prt(self, k, lid, fluxswitch, tide=0.0):
        """The reference's DATA_PRT (output.f90:758-851): writes z_/m_/n_
        files of layer lid at step k into the working directory."""
        self.lib().ref_data_prt(_i(k), _i(lid), _i(fluxswitch), _f(tide))

    def okada(self, params, fswitch=0):
        """The reference's DEFORM_OKADA on layer 1; returns its DEFORM."""
        pa = np.ascontiguousarray(params, np.float32)
        self.lib().ref_okada(_p(pa), _i(fswitch))
        return self.deform(1)[0]

    def floor_deform(self, params, fswitch=0):
        """The reference's GET_FLOOR_DEFORM for one fault rupturing at T = 0."""
        pa = np.ascontiguousarray(params, np.float32)
        big_stack(self.lib().ref_floor_deform, _p(pa), _i(fswitch))

    def sponge(self, on=1):
        """SPONGE_COEF on layer 1 (returns COEFX, COEFY as (ny, nx)); on != 0
        also applies SPONGE_LAYER after every MOMENT of step()."""
        s = self.specs[1]
        cx = np.empty(s.nx * s.ny, np.float32)
        cy = np.empty(s.nx * s.ny, np.float32)
        self.lib().ref_sponge(_i(on), _p(cx), _p(cy))
        return cx.reshape(s.ny, s.nx), cy.reshape(s.ny, s.nx)

    def sc_child(self, laygov, rel_size, rel_time, extent, dt, fric_switch=0, fric_coef=0.013):
        """A Cartesian child of the spherical layer 1 laid out by the reference's
        SUBGRID_MATCHING with SC_OPTION = 1 over extent (X_START, X_END,
        Y_START, Y_END) in degrees; returns its geometry, POS and CXY."""
        ext = np.ascontiguousarray(extent, np.float32)
        nx, ny = ctypes.c_int(0), ctypes.c_int(0)
        dx, dy = ctypes.c_float(0), ctypes.c_float(0)
        corners = np.zeros(4, np.int32)
        self.lib().ref_sc_child(_i(laygov), _i(rel_size), _i(rel_time), _p(ext), _f(dt),
                                _i(fric_switch), _f(fric_coef), ctypes.byref(nx), ctypes.byref(ny),
                                ctypes.byref(dx), ctypes.byref(dy), _p(corners))
        lid = max(self.specs) + 1
        n = nx.value * ny.value
        pos = np.empty(2 * n, np.int32)
        cxy = np.empty(4 * n, np.float32)
        x = np.empty(nx.value, np.float32)
        y = np.empty(ny.value, np.float32)
        self.lib().ref_sc_io(_i(lid), _p(pos), _p(cxy), _p(x), _p(y))
        geom = dict(id=lid, nx=nx.value, ny=ny.value, dx=dx.value, dy=dy.value,
                    corners=tuple(int(c) for c in corners), pos=pos, cxy=cxy, x=x, y=y)
        self.specs[lid] = type("Spec", (), dict(nx=nx.value, ny=ny.value))()
        return geom

    def set_hz(self, lid, h, z0):
        """Depth and initial elevation of a layer (then PQ_DEPTH)."""
        ha = np.ascontiguousarray(h, np.float32)
        za = np.ascontiguousarray(z0, np.float32)
        self.lib().ref_set_hz(_i(lid), _p(ha), _p(za))

    def bc_wall(self):
        """BC_WALL on layer 1 (H of the two outermost rows/columns = -999)."""
        self.lib().ref_bc_wall()

    def deform(self, lid):
        """(DEFORM, H) of layer lid as (ny, nx) REAL arrays."""
        s = self.specs[lid]
        d = np.empty(s.nx * s.ny, np.float32)
        h = np.empty(s.nx * s.ny, np.float32)
        self.lib().ref_deform_io(_i(lid), _p(d), _p(h))
        return d.reshape(s.ny, s.nx), h.reshape(s.ny, s.nx)

    def hot_start(self, start_type, start_step):
        """The reference's HOT_START (hotstart.f90:3-25) in the working
        directory: START_TYPE 0 writes every layer's snapshot files, 1 loads
        layer 1, 2 every layer."""
        self.lib().ref_hot_start(_i(start_type), _i(start_step))

    def zmaxmin(self, lid, zmax=None, zmin=None):
        """Z_MAX and Z_MIN of layer lid: set them when given, else return them."""
        s = self.specs[lid]
        if zmax is not None:
            a = np.ascontiguousarray(zmax, np.float32).reshape(-1)
            b = np.ascontiguousarray(zmin, np.float32).reshape(-1)
            self.lib().ref_zmaxmin_io(_i(lid), _i(1), _p(a), _p(b))
            return None
        a = np.empty(s.nx * s.ny, np.float32)
        b = np.empty(s.nx * s.ny, np.float32)
        self.lib().ref_zmaxmin_io(_i(lid), _i(0), _p(a), _p(b))
        return a.reshape(s.ny, s.nx), b.reshape(s.ny, s.nx)

    def max_amp(self, time, tend, tide=0.0):
        """The reference's MAX_AMP (output.f90:855-1031) at model time
        `time` (REAL): writes zmax files into the working directory when due."""
        self.lib().ref_max_amp(_f(time), _f(tend), _f(tide))
