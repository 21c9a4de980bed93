"""Host mirror of the COMCOT driver's layer loop over the HIP C ABI.

DeviceRun registers a Scenario's layers with gcomcot_init_layer_ex_ (or,
for spherical linear layers, the reference entry gcomcot_init_layer_ with
full PARA planes) and drives the step either as the whole-run device loop
(gcomcot_run_steps_) or call by call through the reference entry points in
the order of the Fortran main loop (comcot.f90:359-439), which is what the
unmodified Fortran driver does when linked against libgcomcot_hip.so.
"""
import ctypes
import os

import numpy as np

from . import scenarios
from ._lib import GcomcotError, arr, check, fp, ip, last_error, lib


def _plane(ny, nx):
    return np.empty((ny, nx), dtype=np.float32)


def _ok(what):
    """Raise if the last C ABI call recorded an error (status mode)."""
    code, msg = last_error()
    if code:
        raise GcomcotError(f"{what}: {msg} (code {code})")


def para_planes(spec):
    """The (NX,NY) planes PARA hands to gcomcot_init_layer_ (R1..R5, R6, R11)."""
    c = scenarios.para_rows(spec.laygov, spec.dt, spec.dx, spec.y, spec.dely).reshape(10, spec.ny)
    h = spec.h
    gx = np.float32(1.0e-5)
    hp = np.zeros_like(h)
    hq = np.zeros_like(h)
    wet = h > gx
    hn = np.concatenate([h[:, 1:], h[:, -1:]], axis=1)   # IP1 clamped at NX
    hu = np.concatenate([h[1:, :], h[-1:, :]], axis=0)   # JP1 clamped at NY
    hp[wet] = (np.float32(0.5) * (h + hn))[wet]
    hq[wet] = (np.float32(0.5) * (h + hu))[wet]
    row = lambda k: np.repeat(c[k][:, None], spec.nx, axis=1).astype(np.float32)
    r2 = (row(3) * np.maximum(hp, np.float32(0))).astype(np.float32)
    r4 = (row(7) * np.maximum(hq, np.float32(0))).astype(np.float32)
    return dict(R1=row(1), R2=r2, R3=row(6), R4=r4, R5=row(8), R6=c[9].astype(np.float32),
                R11=row(2))


class DeviceRun:
    """One scenario resident on the GPU, driven through the C ABI."""

    def __init__(self, scenario, reference_init=False, arrival_threshold=0.01):
        self.scenario = scenario
        L = lib()
        L.gcomcot_set_error_mode_(ip(0))   # errors raise here instead of exiting
        L.cuda_shutdown_()
        if reference_init:  # the reference's start-up call (prints the device)
            L.gcomcot_init_gpu_()
            _ok("gcomcot_init_gpu_")
        L.gcomcot_set_arrival_threshold_(fp(arrival_threshold))
        self.specs = {}
        for s in scenario.layers:
            self.specs[s.id] = s
            corners = np.array(s.corners, dtype=np.int32)
            ha, hp = arr(s.h)
            za, zp = arr(s.z0)
            if reference_init:
                if s.laycord != 0 or s.laygov != 0:
                    raise GcomcotError("the reference init entry only takes spherical linear layers")
                P = para_planes(s)
                keep = {k: arr(v) for k, v in P.items()}
                L.gcomcot_init_layer_(ip(s.id), ip(s.parent), ip(s.level), ip(s.rel_size),
                                      ip(s.rel_time), corners.ctypes.data_as(ctypes.c_void_p),
                                      fp(0), fp(0), keep["R1"][1], keep["R2"][1], keep["R3"][1],
                                      keep["R4"][1], keep["R5"][1], keep["R6"][1],
                                      keep["R11"][1], hp, zp, ip(s.nx), ip(s.ny))
                _ok(f"gcomcot_init_layer_({s.id})")
                continue
            if s.sc_pos is not None:  # SC_OPTION = 1 child
                pa = np.ascontiguousarray(s.sc_pos, np.int32)
                ca = np.ascontiguousarray(s.sc_cxy, np.float32)
                rc = L.gcomcot_init_layer_sc_(ip(s.id), ip(s.parent), ip(s.level), ip(s.laygov),
                                              ip(s.rel_size), ip(s.rel_time),
                                              corners.ctypes.data_as(ctypes.c_void_p),
                                              ip(s.fric_switch), fp(s.fric_coef), fp(s.dt),
                                              fp(s.dx), fp(s.dy), hp, zp, ip(s.nx), ip(s.ny),
                                              pa.ctypes.data_as(ctypes.c_void_p),
                                              ca.ctypes.data_as(ctypes.c_void_p))
                check(rc, f"init SC layer {s.id}")
                continue
            if s.laycord == 0:
                coef = scenarios.para_rows(s.laygov, s.dt, s.dx, s.y, s.dely)
                cptr = coef.ctypes.data_as(ctypes.c_void_p)
            else:
                coef = None
                cptr = None
            rc = L.gcomcot_init_layer_ex_(ip(s.id), ip(s.parent), ip(s.level), ip(s.laycord),
                                          ip(s.laygov), ip(s.rel_size), ip(s.rel_time),
                                          corners.ctypes.data_as(ctypes.c_void_p),
                                          ip(s.fric_switch), fp(s.fric_coef), fp(s.dt),
                                          fp(s.dx), fp(s.dy), cptr, hp, zp, ip(s.nx), ip(s.ny))
            check(rc, f"init layer {s.id}")

    def autotune(self, reps=5):
        """Pick the fastest step kernel shape (library autotuner): for a
        single layer its tile variant or two-step shape, for a nest the
        tile variant of each layer that runs fused steps; returns (tile
        variant, two-step mode or 0, ms per root step of the tuned
        launches)."""
        v, b, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0)
        check(lib().gcomcot_autotune_(ip(reps), ctypes.byref(v), ctypes.byref(b),
                                      ctypes.byref(ms)), "autotune")
        return v.value, b.value, ms.value

    def set_store_policy(self, mode=1):
        """How step launches store the new time level (process-wide): 1 =
        streamed past the caches for layers over ~5.6 M or under ~0.93 M
        cells, 0 = always through the caches, 2 = always streamed; same
        bits."""
        check(lib().gcomcot_set_store_policy_(ip(mode)), "store policy")

    def set_fused_halves(self, on=True):
        """For a linear outermost layer with nested layers: the momentum
        launch of each root step but the last of a run also computes the
        next root step's continuity (on, the default) or not; same bits."""
        check(lib().gcomcot_set_fused_halves_(ip(1 if on else 0)), "fused halves")

    # ------------------------------------------------------------ sources
    def set_coords(self, lids=None):
        """Give the library each layer's X (NX) / Y (NY) coordinates."""
        L = lib()
        for lid in (lids or sorted(self.specs)):
            xa, xp = arr(self.coords(lid)[0])
            ya, yp = arr(self.coords(lid)[1])
            check(L.gcomcot_set_layer_coords_(ip(lid), xp, yp), f"coords of layer {lid}")

    def fault(self, params, switch=0):
        """Apply one fault plane's seafloor deformation on the device; params
        are the FAULT record's 11 values HH, L, W, D, TH, DL, RD, YO, XO, Y0,
        X0 (gcomcot.source.Fault.params())."""
        pa = np.ascontiguousarray(params, np.float32)
        if pa.size != 11:
            raise GcomcotError("a fault has 11 parameters")
        check(lib().gcomcot_fault_deform_(pa.ctypes.data_as(ctypes.c_void_p), ip(switch)),
              "fault deform")

    def save(self, path):
        """Exact checkpoint of the device state to a file."""
        b = os.fsencode(os.fspath(path))
        check(lib().gcomcot_checkpoint_save_(ctypes.create_string_buffer(b), ip(len(b))),
              "checkpoint save")

    def load(self, path):
        """Restore a checkpoint made from a run with the same layers."""
        b = os.fsencode(os.fspath(path))
        check(lib().gcomcot_checkpoint_load_(ctypes.create_string_buffer(b), ip(len(b))),
              "checkpoint load")

    def sponge(self, on=True):
        """SPONGE_LAYER on layer 1 after every momentum pass (coefficients from
        its depth and coordinates; call set_coords first)."""
        check(lib().gcomcot_sponge_set_(ip(1 if on else 0)), "sponge")

    def schedule_fault(self, params, switch=0, t0=0.0):
        """A fault that ruptures at model time t0 during run(): applied where
        the reference's loop calls GET_FLOOR_DEFORM with TIME <= t0 < TIME + DT."""
        pa = np.ascontiguousarray(params, np.float32)
        if pa.size != 11:
            raise GcomcotError("a fault has 11 parameters")
        check(lib().gcomcot_fault_schedule_(pa.ctypes.data_as(ctypes.c_void_p), ip(switch),
                                            fp(t0)), "fault schedule")

    def pending_faults(self):
        return int(lib().gcomcot_fault_pending_())

    def deform(self, lid=1):
        s = self.specs[lid]
        out = _plane(s.ny, s.nx)
        check(lib().gcomcot_get_deform_(ip(lid), out.ctypes.data_as(ctypes.c_void_p)), "deform")
        return out

    def depth(self, lid=1):
        s = self.specs[lid]
        out = _plane(s.ny, s.nx)
        check(lib().gcomcot_get_depth_(ip(lid), out.ctypes.data_as(ctypes.c_void_p)), "depth")
        return out

    # ------------------------------------------------------------ outputs
    def set_output(self, directory, iprt, out_opt=2, tend=0.0, tide=0.0, fluxswitch=None):
        """Reference-format files written by the device loop into `directory`:
        snapshots every `iprt` root steps (and at step 0), the maximum files
        of layer 1 when out_opt is 0 or 2, with TEND `tend` and TIDE_LEVEL
        `tide`; fluxswitch maps layer id -> FLUXSWITCH (default 1, z only)."""
        L = lib()
        for lid, fs in (fluxswitch or {}).items():
            check(L.gcomcot_output_layer_(ip(lid), ip(fs)), f"output layer {lid}")
        d = os.fsencode(os.fspath(directory))
        buf = ctypes.create_string_buffer(d)
        check(L.gcomcot_output_set_(buf, ip(len(d)), ip(iprt), ip(out_opt), fp(tend), fp(tide)),
              "output set")

    def flush_output(self):
        """Wait for every queued file; returns the number written so far."""
        n = ctypes.c_longlong(0)
        check(lib().gcomcot_output_flush_(ctypes.byref(n)), "output flush")
        return n.value

    # ------------------------------------------------------------ running
    def run(self, steps, graph=True):
        rc = lib().gcomcot_run_steps_(ip(steps), ip(1 if graph else 0))
        check(rc, "run_steps")

    def timed(self, steps, per_step=True):
        total = ctypes.c_float(0)
        each = np.zeros(steps, dtype=np.float32) if per_step else None
        rc = lib().gcomcot_time_steps_(ip(steps), ctypes.byref(total),
                                       each.ctypes.data_as(ctypes.c_void_p) if per_step else None)
        check(rc, "time_steps")
        return total.value, each

    def dropin_step(self):
        """One root step through the reference entry points, in the order
        of comcot.f90:362-439 (MAX_AMP after momentum, CHANGE last)."""
        L = lib()
        one = ctypes.c_int(1)
        L.mass_launch_(None, None, None, ctypes.byref(one))
        L.openbd_launch_(None)
        if len(self.scenario.layers) > 1:
            L.all_grid_launch_()
        L.momt_launch_(None, None, None, ctypes.byref(one))
        L.maxamp_launch_(ctypes.byref(one))
        L.cuda_update_layer_(ctypes.byref(one))
        _ok("drop-in step")

    # ------------------------------------------------------------ gauges
    def set_gauges(self, points, capacity):
        """Gauges [(x0, y0, layer id), ...] sampled on the device after every
        root step; each layer's coordinates are its LayerSpec lon/lat grid
        (`coords`)."""
        L = lib()
        for lid in sorted({p[2] for p in points}):
            x, y = self.coords(lid)
            xa, xp = arr(x)
            ya, yp = arr(y)
            check(L.gcomcot_set_layer_coords_(ip(lid), xp, yp), f"coords of layer {lid}")
        x0 = np.array([p[0] for p in points], np.float32)
        y0 = np.array([p[1] for p in points], np.float32)
        ids = np.array([p[2] for p in points], np.int32)
        check(L.gcomcot_gauges_set_(ip(len(points)), x0.ctypes.data_as(ctypes.c_void_p),
                                    y0.ctypes.data_as(ctypes.c_void_p),
                                    ids.ctypes.data_as(ctypes.c_void_p), ip(capacity)),
              "set gauges")
        self.ngauges = len(points)

    def gauges(self, max_steps):
        out = np.zeros((self.ngauges, max_steps), np.float32)
        n = ctypes.c_int(0)
        check(lib().gcomcot_gauges_read_(out.ctypes.data_as(ctypes.c_void_p), ip(max_steps),
                                         ctypes.byref(n)), "read gauges")
        return out[:, :n.value]

    def coords(self, lid):
        """X (NX) and Y (NY) of a layer as float32: its longitudes/latitudes
        (spherical) or metre offsets (Cartesian) from the scenario."""
        s = self.specs[lid]
        return scenarios.layer_coords(s)

    # ------------------------------------------------------------ state
    def state(self, lid=1):
        """Level-1 and level-2 planes of a layer, as (ny, nx) float32."""
        s = self.specs[lid]
        L = lib()
        out = {k: _plane(s.ny, s.nx) for k in ("z1", "z2", "m1", "m2", "n1", "n2")}
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        L.cuda_getz1_(p(out["z1"]), ip(lid))
        L.cuda_getz_(p(out["z2"]), ip(lid))
        L.cuda_getmn1_(p(out["m1"]), p(out["n1"]), ip(lid))
        L.cuda_getmn_(p(out["m2"]), p(out["n2"]), ip(lid))
        _ok("state")
        return out

    def set_level1(self, lid, z=None, m=None, n=None):
        s = self.specs[lid]
        L = lib()
        if z is not None:
            za, zp = arr(z)
            L.cuda_setz1_(zp, ip(lid))
        if m is not None or n is not None:
            cur = self.state(lid)
            ma, mp = arr(cur["m1"] if m is None else m)
            na, np_ = arr(cur["n1"] if n is None else n)
            L.cuda_setmn1_(mp, np_, ip(lid))

    def maxima(self, lid=1):
        s = self.specs[lid]
        zw = _plane(s.ny, s.nx)
        ar = np.empty((s.ny, s.nx), dtype=np.int32)
        zr = _plane(s.ny, s.nx)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        lib().gcomcot_get_max_(ip(lid), p(zw), p(ar))
        lib().cuda_getzmax_(p(zr), ip(lid))
        return {"zmax_wet": zw, "arrival": ar, "zmax_raw": zr}

    def flux_cache(self, lid=1):
        s = self.specs[lid]
        xf = np.empty(2 * s.nx, dtype=np.float32)
        yf = np.empty(2 * s.ny, dtype=np.float32)
        lib().gcomcot_get_flux_cache_(ip(lid), xf.ctypes.data_as(ctypes.c_void_p),
                                      yf.ctypes.data_as(ctypes.c_void_p))
        return xf.reshape(2, s.nx), yf.reshape(2, s.ny)

    def close(self):
        lib().cuda_shutdown_()
