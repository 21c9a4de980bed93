"""Host side of the scenario batch engine (include/gcomcot_hip.h part 3).

Batch registers Scenarios (a root layer and at most one level of nested
layers, each with its own root-step count) with the library, advances
them together and reads back their planes, maxima and arrival fields.
The engine itself (plans, persistent step kernel, HIP graph replay,
tile-shape autotuner) is native code in csrc/batch.hip; this class only
marshals arrays.
"""
import ctypes
import os

import numpy as np

from . import scenarios
from ._lib import GcomcotError, arr, check, fp, ip, lib

_LL = ctypes.c_longlong


def _ll(v):
    return ctypes.byref(_LL(int(v)))


def _setup(L):
    if getattr(L, "_batch_typed", False):
        return
    for name in ("gcomcot_batch_device_bytes_", "gcomcot_layer_footprint_"):
        getattr(L, name).restype = _LL
    L._batch_typed = True


def layer_footprint(nx, ny, nonlinear):
    """Device bytes one layer takes (planes, tables; 256-B aligned)."""
    L = lib()
    _setup(L)
    return int(L.gcomcot_layer_footprint_(ip(nx), ip(ny), ip(1 if nonlinear else 0)))


def scenario_footprint(sc):
    return sum(layer_footprint(s.nx, s.ny, _is_nl(s)) for s in sc.layers)


def _is_nl(s):
    if s.laycord == 0:
        return s.laygov in (1, 9)
    return s.laygov in (1, 5, 9)


class Batch:
    """The library's batch engine (one per process)."""

    def __init__(self, arrival_threshold=0.01, autotune=True, shape=-1, graphs=True,
                 two_step=True, two_step_shape=None, streams=3, fused_halves=True):
        """two_step: scenarios whose only layer is linear may advance two
        root steps per launch, in tile shape two_step_shape (the library's
        two-step shapes 0-6), or, when None, in the shape the autotuner
        finds fastest if that beats two one-step launches.  streams: 3 =
        single-layer scenarios and two chains of nested scenarios side by
        side, 2 = single-layer scenarios beside one chain, 1 = one stream.
        fused_halves: the momentum launch of a nested linear root also
        computes its next continuity (same bits, one launch fewer per step)."""
        L = lib()
        _setup(L)
        L.gcomcot_set_error_mode_(ip(0))
        check(L.gcomcot_batch_reset_(), "batch reset")
        L.gcomcot_batch_set_arrival_threshold_(fp(arrival_threshold))
        check(L.gcomcot_batch_set_tuning_(ip(1 if autotune else 0), ip(shape)), "batch tuning")
        check(L.gcomcot_batch_set_graphs_(ip(1 if graphs else 0)), "batch graphs")
        check(L.gcomcot_batch_set_two_step_(ip(1 if two_step else 0),
                                            None if two_step_shape is None else ip(two_step_shape)),
              "batch two-step")
        check(L.gcomcot_batch_set_streams_(ip(streams)), "batch streams")
        check(L.gcomcot_batch_set_fused_halves_(ip(1 if fused_halves else 0)),
              "batch fused halves")
        self.specs = {}

    def add(self, sc, steps=None):
        """Register scenario `sc`; returns its id.  steps defaults to sc.steps."""
        L = lib()
        sid = ctypes.c_int(0)
        check(L.gcomcot_batch_new_scenario_(_ll(sc.steps if steps is None else steps),
                                            ctypes.byref(sid)), "batch new scenario")
        try:
            self._add_layers(sid, sc)
        except GcomcotError:
            lib().gcomcot_batch_remove_(ctypes.byref(sid))  # no half-registered scenario
            raise
        self.specs[sid.value] = sc
        return sid.value

    def _add_layers(self, sid, sc):
        L = lib()
        keep = []
        for s in sc.layers:
            corners = np.array(s.corners, dtype=np.int32)
            ha, hp = arr(s.h)
            za, zp = arr(s.z0)
            keep += [corners, ha, za]
            if getattr(s, "sc_pos", None) is not None:  # SC_OPTION = 1 child
                pa = np.ascontiguousarray(s.sc_pos, np.int32)
                ca = np.ascontiguousarray(s.sc_cxy, np.float32)
                keep += [pa, ca]
                rc = L.gcomcot_batch_add_layer_sc_(
                    ctypes.byref(sid), ip(s.id), ip(s.parent), ip(s.laygov), ip(s.rel_size),
                    ip(s.rel_time), corners.ctypes.data_as(ctypes.c_void_p), ip(s.fric_switch),
                    fp(s.fric_coef), fp(s.dt), fp(s.dx), fp(s.dy), hp, zp, ip(s.nx), ip(s.ny),
                    pa.ctypes.data_as(ctypes.c_void_p), ca.ctypes.data_as(ctypes.c_void_p))
                check(rc, f"batch add SC layer {s.id} of {sc.name}")
                continue
            if s.laycord == 0:
                coef = scenarios.para_rows(s.laygov, s.dt, s.dx, s.y, s.dely)
                keep.append(coef)
                cptr = coef.ctypes.data_as(ctypes.c_void_p)
            else:
                cptr = None
            rc = L.gcomcot_batch_add_layer_(
                ctypes.byref(sid), ip(s.id), ip(s.parent), ip(s.laycord), ip(s.laygov),
                ip(s.rel_size), ip(s.rel_time), corners.ctypes.data_as(ctypes.c_void_p),
                ip(s.fric_switch), fp(s.fric_coef), fp(s.dt), fp(s.dx), fp(s.dy), cptr, hp, zp,
                ip(s.nx), ip(s.ny))
            check(rc, f"batch add layer {s.id} of {sc.name}")

    def save(self, sid, path):
        """Exact checkpoint of scenario sid (its planes and step count)."""
        b = os.fsencode(os.fspath(path))
        check(lib().gcomcot_batch_checkpoint_save_(ip(sid), ctypes.create_string_buffer(b),
                                                   ip(len(b))), "batch checkpoint save")

    def load(self, sid, path):
        """Resume scenario sid (registered with the same layers) from a checkpoint."""
        b = os.fsencode(os.fspath(path))
        check(lib().gcomcot_batch_checkpoint_load_(ip(sid), ctypes.create_string_buffer(b),
                                                   ip(len(b))), "batch checkpoint load")

    def fault(self, sid, params, switch=0):
        """Seafloor deformation of one fault plane for scenario sid, applied on
        the device (spherical layers get their coordinates from the spec)."""
        L = lib()
        for s in self.specs[sid].layers:
            x, y = scenarios.layer_coords(s)
            xa, xp = arr(x)
            ya, yp = arr(y)
            check(L.gcomcot_batch_set_coords_(ip(sid), ip(s.id), xp, yp), "batch coords")
        pa = np.ascontiguousarray(params, np.float32)
        check(L.gcomcot_batch_fault_deform_(ip(sid), pa.ctypes.data_as(ctypes.c_void_p),
                                            ip(switch)), "batch fault")

    def run(self, nsteps):
        """Advance every unfinished scenario by up to nsteps; returns #finished."""
        nf = ctypes.c_int(0)
        check(lib().gcomcot_batch_run_(_ll(nsteps), ctypes.byref(nf)), "batch run")
        return nf.value

    def finished(self):
        L = lib()
        cap = max(1, len(self.specs))
        buf = (ctypes.c_int * cap)()
        n = L.gcomcot_batch_finished_(buf, ip(cap))
        return [buf[i] for i in range(min(n, cap))]

    def progress(self, sid):
        d, t = _LL(0), _LL(0)
        check(lib().gcomcot_batch_progress_(ip(sid), ctypes.byref(d), ctypes.byref(t)),
              "batch progress")
        return d.value, t.value

    def _spec(self, sid, lid):
        for s in self.specs[sid].layers:
            if s.id == lid:
                return s
        raise GcomcotError(f"scenario {sid} has no layer {lid}")

    def state(self, sid, lid=1):
        s = self._spec(sid, lid)
        out = {k: np.empty((s.ny, s.nx), np.float32) for k in ("z1", "m1", "n1")}
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(lib().gcomcot_batch_get_state_(ip(sid), ip(lid), p(out["z1"]), p(out["m1"]),
                                             p(out["n1"])), "batch state")
        return out

    def maxima(self, sid, lid=1):
        s = self._spec(sid, lid)
        zw = np.empty((s.ny, s.nx), np.float32)
        ar = np.empty((s.ny, s.nx), np.int32)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        check(lib().gcomcot_batch_get_max_(ip(sid), ip(lid), p(zw), p(ar)), "batch max")
        return {"zmax_wet": zw, "arrival": ar}

    def remove(self, sid):
        check(lib().gcomcot_batch_remove_(ip(sid)), "batch remove")
        self.specs.pop(sid, None)

    def device_bytes(self, sid=0):
        return int(lib().gcomcot_batch_device_bytes_(ip(sid)))

    def info(self):
        v = [_LL(0) for _ in range(4)]
        t = ctypes.c_int(0)
        lib().gcomcot_batch_info_(*[ctypes.byref(x) for x in v], ctypes.byref(t))
        return {"launches": v[0].value, "graph_replays": v[1].value, "plans": v[2].value,
                "cells_per_step": v[3].value, "tuned": t.value}

    def close(self):
        lib().gcomcot_batch_reset_()
        self.specs.clear()
