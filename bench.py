"""Benchmark of the COMCOT shallow-water step on one MI355X.

Workload (default): the single-layer nonlinear spherical configuration,
4096 x 3072 cells at 0.5 arc-min, Manning friction n = 0.013, moving
shoreline, wet-only maximum and first-arrival fields accumulated on the
device, 5000 steps, FP32 — one kernel launch per time step.

Timed region: K steps of the device loop with every input resident in
HBM, bracketed by device synchronisation.  HIP events recorded on the
library's stream between steps give each step's duration; for a single
layer each step is one launch of the fused step kernel, so their mean is
that kernel's average duration, from which the achieved algorithmic
bandwidth is computed (36 B per cell update: Z, M, N, H read, Z, M, N
written, the maximum read and written).

CPU baseline: the reference's own Fortran (flang-built from its sources
into oracle/_ref) timed on a bounded sample (the first steps of the same
workload, one core); the C++ restatement stands in when that build is
absent.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-comcot_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"linear": 28, "nonlinear_max": 36}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=3072)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="steps of the bounded CPU-baseline sample (0 = skip)")
    ap.add_argument("--tile-variant", type=int, default=-1)
    return ap.parse_args()


def cpu_baseline(sc, nsteps):
    """Time the reference Fortran (or the restatement) on nsteps of sc."""
    from oracle import bind
    if bind.Reference.available():
        kind = "reference"
        runner = bind.Reference(sc)
    else:
        kind = "port"
        runner = bind.Oracle(sc, 32, pow_mode=0)
    runner.step(1)  # first step outside the clock (page-in)
    t0 = time.perf_counter()
    runner.step(nsteps)
    dt = time.perf_counter() - t0
    cells = sc.cell_updates_per_step()
    return {"value": cells * nsteps / dt / 1e9, "unit": "GCUPS", "cores": 1, "kind": kind,
            "sample": f"{nsteps} steps (after 1 untimed) of the same "
                      f"{sc.layers[0].nx}x{sc.layers[0].ny} workload, single thread, "
                      f"{dt:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {a.gpus}); scenario-level "
              "sharding across GPUs is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    from gcomcot import DeviceRun, config1
    from gcomcot._lib import ip, lib

    sc = config1(nx=a.nx, ny=a.ny, steps=a.steps)
    cells = sc.cell_updates_per_step()
    if a.tile_variant >= 0:
        lib().gcomcot_set_tile_variant_(ip(a.tile_variant))
    run = DeviceRun(sc)
    run.run(a.warmup, graph=False)
    lib().gcomcot_run_steps_(ip(0), ip(0))  # sync
    t0 = time.perf_counter()
    ms_total, each = run.timed(a.steps)
    t1 = time.perf_counter()
    wall = t1 - t0
    # the first step's interval also holds the queue start; use the rest
    per = each[1:] if len(each) > 1 else each
    kern_ms = float(np.mean(per))
    bpc = BYTES_PER_CELL["nonlinear_max"]
    achieved = bpc * cells / (kern_ms * 1e-3) / 1e9
    gcups = cells * a.steps / wall / 1e9
    mx = run.maxima(1)
    arrived = int(np.count_nonzero(mx["arrival"] >= 0))
    run.close()

    cpu = None
    if a.cpu_steps > 0:
        csc = config1(nx=a.nx, ny=a.ny, steps=a.cpu_steps)
        cpu = cpu_baseline(csc, a.cpu_steps)

    line = {
        "metric": "GCUPS + %HBM roofline; batched scenarios/s with p50/p99 time-to-forecast",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded bathymetry, dipole source)",
        "config": {"workload": f"single-layer nonlinear spherical {a.nx}x{a.ny} at 0.5 arc-min, "
                               "Manning 0.013, wet max + arrival on device",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "kernel_ms_mean": round(kern_ms, 5),
        "kernel_ms_p50": round(float(np.median(per)), 5),
        "event_ms_total": round(ms_total, 3),
        "cells_arrived": arrived,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
