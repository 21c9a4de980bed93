"""Benchmark of the COMCOT shallow-water step on one MI355X.

Headline (value): the single-layer nonlinear spherical configuration,
4096 x 3072 cells at 0.5 arc-min, Manning friction n = 0.013, moving
shoreline, wet-only maximum and first-arrival fields accumulated on the
device, 5000 steps, FP32 -- one launch of the fused step kernel per time
step.  value = cell updates / wall time of the timed steps (GCUPS).

Timed region: K steps with every input resident in HBM, bracketed by
device synchronisation.  HIP events recorded on the library's stream
between steps give each step's duration; for a single layer each step is
one launch of the fused step kernel, so their mean is that kernel's
average duration, from which the achieved algorithmic bandwidth is
computed (36 B per cell update: Z, M, N, H read, Z, M, N written, the
maximum read and written).  roofline.traffic is the HBM bytes per launch
measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE, each scaled by the
same counter's reading for a copy of known size), read from
profiles/traffic.json when it holds an entry for this workload.

CPU baseline: the reference's own Fortran (flang-built from its sources
into oracle/_ref) timed on a bounded sample (the first steps of the same
workload, one core); the C++ restatement stands in when that build is
absent.

Batch (batch.*): the 64-scenario mix (parents 256x256..2048x1536 at 1-4
arc-min, a quarter with a nonlinear Cartesian child, 1000-4000 steps each)
run through the batch engine, against the same scenarios run one at a
time through the whole-run device loop, and against the reference's
per-step call sequence (drop-in entry points, synchronous per call)
timed on a sample of steps.  Serving (serve.*): Poisson request streams
(200 requests at each rate) drawn from the same generator plus 10 %
4096x3072 nonlinear cases with fresh sources, through the
continuous-batching loop; p50/p99 time-to-forecast and scenarios/s over the
whole stream and in steady state, and whether the offered rate is
sustained.  The FP64 oracle's error on sampled served cases is part of the
CPU-baseline leg (cpu_baseline.served_fp64).

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
METRIC = "GCUPS + %HBM roofline; batched scenarios/s with p50/p99 time-to-forecast"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=3072)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="steps of the bounded CPU-baseline sample (0 = skip)")
    ap.add_argument("--tile-variant", type=int, default=-1,
                    help="tile variant of the step kernel (-1: autotune)")
    ap.add_argument("--others", type=int, default=1,
                    help="also time the linear single layer and the two-level nest (0 = skip)")
    ap.add_argument("--batch", type=int, default=1, help="run the 64-scenario batch (0 = skip)")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-two-step", type=int, default=1,
                    help="single-layer linear roots take two root steps per launch (0 = one)")
    ap.add_argument("--batch-streams", type=int, default=3,
                    help="streams of a batch plan (1-3; 1 makes kernel times additive in a trace)")
    ap.add_argument("--dropin-sample", type=int, default=100,
                    help="steps per scenario of the drop-in call-sequence sample")
    ap.add_argument("--serve", type=int, default=1, help="run the serving streams (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="5,10,20")
    ap.add_argument("--serve-pool", type=int, default=120,
                    help="distinct small cases the serving streams draw from")
    ap.add_argument("--serve-fp64", type=int, default=2,
                    help="served small cases per rate checked against the FP64 oracle")
    ap.add_argument("--outputs", type=int, default=1,
                    help="time the device loop with reference-format output files (0 = skip)")
    ap.add_argument("--outputs-steps", type=int, default=1000)
    ap.add_argument("--outputs-iprt", type=int, default=250)
    ap.add_argument("--source", type=int, default=1,
                    help="time a fault source on the headline grid (0 = skip)")
    ap.add_argument("--accuracy", type=int, default=1,
                    help="with the CPU baseline: FP32 vs FP64-oracle error on a reduced grid")
    ap.add_argument("--graphs", type=int, default=1,
                    help="batch engine replays HIP graphs (0: enqueue each step)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the streaming-copy probe on MIB MiB (counter calibration)")
    ap.add_argument("--linear-probe", default="", metavar="NXxNY",
                    help="only run the linear single layer of that size (profiling)")
    ap.add_argument("--blocking", type=int, default=4,
                    help="with --linear-probe: 0 one step per launch, 1-7 two-step shapes")
    return ap.parse_args()


def copy_probe(mib, reps=20):
    import ctypes
    from gcomcot._lib import check, fp, ip, lib
    ms = ctypes.c_float(0)
    check(lib().gcomcot_hbm_copy_(ip(mib), ip(reps), ctypes.byref(ms)), "copy probe")
    gbs = 2 * (mib << 20) / (ms.value * 1e-3) / 1e9
    return {"mib": mib, "ms_per_copy": round(ms.value, 5), "read_plus_write_GBs": round(gbs, 1)}


def linear_probe(a):
    """The linear single layer (C0's generator at the given size), --steps
    root steps through the device loop without graphs (one launch per step
    or per two steps, as --blocking says; tile variant --tile-variant when
    set), for kernel traces and counters."""
    from gcomcot import DeviceRun, linear_layer
    from gcomcot._lib import ip, lib
    nx, ny = (int(v) for v in a.linear_probe.lower().split("x"))
    sc = linear_layer(nx=nx, ny=ny, steps=a.steps)
    lib().gcomcot_set_temporal_blocking_(ip(a.blocking))
    if a.tile_variant >= 0:
        lib().gcomcot_set_tile_variant_(ip(a.tile_variant))
    run = DeviceRun(sc)
    run.run(2, graph=False)
    t0 = time.perf_counter()
    run.run(a.steps, graph=False)  # returns once the device is idle
    wall = time.perf_counter() - t0
    run.close()
    return {"workload": sc.name, "blocking": a.blocking, "steps": a.steps,
            "us_per_step": round(wall / a.steps * 1e6, 2),
            "gcups": round(nx * ny * a.steps / wall / 1e9, 2)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def host_cpu():
    """nproc and model name of the host the CPU figures were taken on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_all_cores(sc, nsteps):
    """The FP32 restatement (the reference's libm pow) on every OpenMP
    thread the host gives it, on the same sample: the cell loops are
    Jacobi-style, so its results equal the serial run's.  Reported beside
    the single-core reference, labelled as the restatement."""
    from oracle import bind
    runner = bind.Oracle(sc, 32, pow_mode=0)
    runner.step(1)
    t0 = time.perf_counter()
    runner.step(nsteps)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {"value": sc.cell_updates_per_step() * nsteps / dt / 1e9, "unit": "GCUPS",
            "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps (after 1 untimed), {threads} OpenMP threads, {dt:.1f} s"}


def traffic_entry(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload_key)
    except (OSError, ValueError):
        return None


def bench_c1(a):
    from gcomcot import DeviceRun, nonlinear_layer
    from gcomcot._lib import ip, lib

    log(f"nonlinear layer {a.nx}x{a.ny}: generating")
    sc = nonlinear_layer(nx=a.nx, ny=a.ny, steps=a.steps)
    cells = sc.cell_updates_per_step()
    run = DeviceRun(sc)
    if a.tile_variant >= 0:
        lib().gcomcot_set_tile_variant_(ip(a.tile_variant))
        variant = a.tile_variant
    else:
        variant = run.autotune()[0]  # state-neutral library autotuner
    run.run(a.warmup, graph=False)
    lib().gcomcot_run_steps_(ip(0), ip(0))  # sync
    log(f"nonlinear layer: timing {a.steps} steps (tile variant {variant})")
    t0 = time.perf_counter()
    ms_total, each = run.timed(a.steps)
    wall = time.perf_counter() - t0
    per = each[1:] if len(each) > 1 else each  # the first interval holds the queue start
    kern_ms = float(np.mean(per))
    bpc = BYTES_PER_CELL["nonlinear_max"]
    achieved = bpc * cells / (kern_ms * 1e-3) / 1e9
    gcups = cells * a.steps / wall / 1e9
    mx = run.maxima(1)
    arrived = int(np.count_nonzero(mx["arrival"] >= 0))
    run.close()
    import ctypes
    tx = (ctypes.c_int * 64)()
    ty = (ctypes.c_int * 64)()
    n = ctypes.c_int(0)
    lib().gcomcot_tile_variants_(ctypes.byref(n), tx, ty)
    key = f"nonlinear-{a.nx}x{a.ny}"
    tr = traffic_entry(key)
    traffic = None
    if tr and tr.get("tile") == [tx[variant], ty[variant]]:
        traffic = tr.get("bytes_per_launch")
    return sc, cells, gcups, wall, kern_ms, achieved, traffic, arrived, ms_total, \
        (tx[variant], ty[variant])


def bench_other_configs(a):
    """The single-layer linear configuration (1024x768, 2000 steps) and the
    two-level nest (2048x1536 linear parent + 1801x1801 nonlinear Cartesian
    child, REL_SIZE 4, REL_TIME 4, 3000 steps) through the whole-run device
    loop (HIP graph): GCUPS over all layers' cell updates."""
    from gcomcot import DeviceRun, linear_layer, two_level_nest
    out = {}
    for name, sc, bpc in (("linear_1024x768", linear_layer(), BYTES_PER_CELL["linear"]),
                          ("nest_2048x1536_1801x1801", two_level_nest(), BYTES_PER_CELL["linear"])):
        log(f"{name}: {sc.name}, {sc.steps} steps")
        run = DeviceRun(sc)
        tuned = run.autotune()
        run.run(10, graph=bool(a.graphs))
        t0 = time.perf_counter()
        run.run(sc.steps, graph=bool(a.graphs))
        wall = time.perf_counter() - t0
        cells = sc.cell_updates_per_step()
        run.close()
        gcups = cells * sc.steps / wall / 1e9
        out[name] = {"workload": sc.name, "steps": sc.steps, "cell_updates_per_step": cells,
                     "tuned": {"tile_variant": tuned[0], "two_step_shape": tuned[1]},
                     "gcups": round(gcups, 3), "ms_per_step": round(wall * 1e3 / sc.steps, 5),
                     "algorithmic_GBs": round(gcups * bpc, 1),
                     "frac_of_8TBs": round(gcups * bpc / HBM_PEAK_GBS, 4)}
    return out


def bench_outputs(a):
    """The headline workload through the device loop (graphs), without and
    with the reference's output files: snapshots z/m/n every --outputs-iprt
    root steps and at step 0, MAX_AMP's maximum kept by the step kernel and
    written when due (TEND at the last step).  ms_per_step_* time the loop
    call; files still being written when it returns are waited for after it
    (final_flush_ms).  Files go to a temporary directory, removed afterwards."""
    import shutil
    import tempfile

    from gcomcot import DeviceRun, nonlinear_layer
    n = a.outputs_steps
    sc = nonlinear_layer(nx=a.nx, ny=a.ny, steps=n)
    dt = np.float32(sc.layers[0].dt)
    out = {"steps": n, "iprt": a.outputs_iprt, "fluxswitch": 0, "out_opt": 2}
    walls = {}
    d = tempfile.mkdtemp(prefix="gcomcot_out_")
    try:
        for mode in ("off", "on"):
            log(f"outputs {mode}: {n} steps")
            run = DeviceRun(sc)
            run.autotune()
            if mode == "on":
                run.set_output(d, a.outputs_iprt, 2, np.float32((n - 1) * dt), 0.0, {1: 0})
            t0 = time.perf_counter()
            run.run(n, graph=bool(a.graphs))
            walls[mode] = time.perf_counter() - t0
            if mode == "on":
                t1 = time.perf_counter()
                out["files_written"] = run.flush_output()
                out["final_flush_ms"] = round((time.perf_counter() - t1) * 1e3, 2)
                out["bytes_on_disk"] = sum(os.path.getsize(os.path.join(d, f))
                                           for f in os.listdir(d))
            run.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    out["ms_per_step_off"] = round(walls["off"] * 1e3 / n, 5)
    out["ms_per_step_on"] = round(walls["on"] * 1e3 / n, 5)
    out["overhead_frac"] = round(walls["on"] / walls["off"] - 1.0, 4)
    return out


def bench_source(a):
    """A fault source on the headline grid: the library's device Okada
    (FP64) + spreading + update, wall time of gcomcot_fault_deform_ (second
    call; the first allocates the deformation planes), against the
    reference's own DEFORM_OKADA (flang, REAL, one core) on a 512x384 sample
    of the same kind of layer."""
    from gcomcot import DeviceRun, nonlinear_layer
    from gcomcot import scenarios as S
    from gcomcot.source import Fault
    sc = nonlinear_layer(nx=a.nx, ny=a.ny, steps=1)
    L = sc.layers[0]
    x, y = S.layer_coords(L)
    f = Fault(10000, 200000, 80000, 6, 30, 15, 90, x0=float(x[L.nx // 2]), y0=float(y[L.ny // 2]))
    run = DeviceRun(sc)
    run.set_coords()
    log(f"source: Okada on {L.nx}x{L.ny}")
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        run.fault(f.params(), f.switch)
        walls.append(time.perf_counter() - t0)
    run.close()
    cells = L.nx * L.ny
    ms = min(walls[1:]) * 1e3
    out = {"workload": f"one fault plane on {L.nx}x{L.ny} spherical", "ms": round(ms, 3),
           "cells_per_s": round(cells / (ms * 1e-3), 1), "precision": "f64"}
    try:
        from oracle.bind import Reference
        if Reference.available():
            small = S.custom(laycord=0, laygov=0, nx=512, ny=384, dx=L.dx)
            sx, sy = S.layer_coords(small.layers[0])
            fs = Fault(10000, 200000, 80000, 6, 30, 15, 90, x0=float(sx[256]), y0=float(sy[192]))
            ref = Reference(small)
            ref.set_coords(1, sx, sy)
            t0 = time.perf_counter()
            ref.okada(fs.params(), fs.switch)
            w = time.perf_counter() - t0
            out["cpu_reference"] = {"sample": "DEFORM_OKADA on 512x384, flang -O2, 1 core",
                                    "cells_per_s": round(512 * 384 / w, 1)}
    except Exception as e:  # the CPU leg is a reported baseline, not the product
        out["cpu_reference"] = {"error": str(e)[:200]}
    return out


def bench_batch(a):
    from gcomcot import DeviceRun
    from gcomcot import scenarios as S
    from gcomcot.batch import Batch

    log(f"batch: generating {a.batch_cases} scenarios")
    cs = S.scenario_mix(n=a.batch_cases)
    upd = [sum(L.nx * L.ny * (L.rel_time if L.id > 1 else 1) for L in c.layers) * c.steps
           for c in cs]
    total = float(sum(upd))
    b = Batch(graphs=bool(a.graphs), two_step=bool(a.batch_two_step), streams=a.batch_streams)
    for c in cs:
        b.add(c)
    log("batch: running the batch engine")
    t0 = time.perf_counter()
    b.run(10 ** 9)
    t_batch = time.perf_counter() - t0
    info = b.info()
    b.close()
    log("batch: one scenario at a time, whole-run device loop")
    t_alone = 0.0
    for c in cs:
        d = DeviceRun(c)
        t0 = time.perf_counter()
        d.run(c.steps, graph=bool(a.graphs))
        t_alone += time.perf_counter() - t0
        d.close()
    log("batch: one scenario at a time, reference call sequence (sample)")
    t_dropin = 0.0
    k = max(1, a.dropin_sample)
    for c in cs:
        d = DeviceRun(c)
        d.dropin_step()
        t0 = time.perf_counter()
        for _ in range(k):
            d.dropin_step()
        t_dropin += (time.perf_counter() - t0) / k * c.steps
        d.close()
    n = len(cs)
    return {
        "scenarios": n, "nested": sum(len(c.layers) > 1 for c in cs),
        "cell_updates": total,
        "batched_s": round(t_batch, 3), "batched_scenarios_per_s": round(n / t_batch, 3),
        "batched_gcups": round(total / t_batch / 1e9, 3),
        "alone_loop_s": round(t_alone, 3), "alone_loop_scenarios_per_s": round(n / t_alone, 3),
        "alone_dropin_s_est": round(t_dropin, 3),
        "alone_dropin_scenarios_per_s": round(n / t_dropin, 3),
        "speedup_vs_alone_loop": round(t_alone / t_batch, 3),
        "speedup_vs_dropin": round(t_dropin / t_batch, 3),
        "dropin_sample": f"{k} steps per scenario, scaled to its step count",
        "launches": info["launches"], "graph_replays": info["graph_replays"],
        "plans": info["plans"],
    }


def bench_serve(a):
    """Continuous serving (config 4): Poisson streams of --serve-requests
    requests at each of --serve-rates, drawn from the config-3 generator
    with a 10% share of 4096x3072 nonlinear cases.  Small cases come from
    one pool shared by the rates; each rate gets fresh large cases (new
    sources on the large world).  Returns the per-rate summaries and, for
    the checker leg, the forecasts kept for --serve-fp64 of the smallest
    small cases of each rate."""
    from gcomcot import scenarios as S
    from gcomcot.serve import Request, Server, poisson_arrivals, summarize

    rates = [float(r) for r in a.serve_rates.split(",") if r]
    nreq = a.serve_requests
    nbig = max(1, round(0.1 * nreq))
    log(f"serve: generating a pool of {a.serve_pool} small cases")
    pool = S.scenario_mix(n=a.serve_pool, case=10)
    out, kept = {}, []
    for ri, rate in enumerate(rates):
        log(f"serve: {nbig} fresh large cases for {rate:g} req/s")
        bigs = S.nonlinear_layer_sources(nbig, case=10, seed=ri)
        rng = np.random.default_rng(300 + ri)
        small = [pool[k] for k in rng.integers(0, len(pool), nreq - nbig)]
        cases = small + bigs
        order = rng.permutation(len(cases))
        t = poisson_arrivals(rate, len(cases), seed=100 + ri)
        reqs = [Request(i, t[i], cases[j]) for i, j in enumerate(order)]
        cost = {r.rid: r.scenario.cell_updates_per_step() * r.steps for r in reqs}
        smalls = [r for r, j in zip(reqs, order) if j < len(small)]
        check = {r.rid for r in sorted(smalls, key=lambda r: cost[r.rid])[:a.serve_fp64]}
        log(f"serve: {nreq} requests at {rate:g} req/s")
        srv = Server(chunk=128, keep_results=check)
        done = srv.serve(reqs)
        srv.close()
        s = summarize(done)
        s["rate_req_per_s"] = rate
        s["large_cases"] = nbig
        s["peak_running"] = srv.peak_running
        out[f"{rate:g}"] = _rounded(s)
        kept += [(f"{rate:g}", r) for r in done if r.result is not None]
    return out, kept


def _rounded(d):
    if isinstance(d, dict):
        return {k: _rounded(v) for k, v in d.items()}
    return round(d, 4) if isinstance(d, float) else d


def cpu_served_accuracy(kept):
    """Part of the CPU-baseline leg: the FP64 oracle's wet maximum and
    arrival field for sampled served cases against the forecasts the
    serving loop returned (FP32 on the device); the oracle is the checker."""
    from oracle import bind
    rows = []
    for rate, r in kept:
        sc = r.scenario
        o64 = bind.Oracle(sc, 64, pow_mode=0)
        o64.step(r.steps)
        om = o64.maxima(1)
        g = r.result
        zd = np.abs(np.asarray(g["zmax_wet"], np.float64) - np.asarray(om["zmax_wet"], np.float64))
        rows.append({"rate": rate, "case": sc.name, "cells": int(sc.cell_updates_per_step()),
                     "steps": int(r.steps), "zmax_max_abs_m": float(zd.max()),
                     "zmax_peak_m": float(np.max(om["zmax_wet"])),
                     "arrival_cells": int(np.count_nonzero(om["arrival"] >= 0)),
                     "arrival_mismatch": int(np.count_nonzero(g["arrival"] != om["arrival"]))})
    return rows


def cpu_accuracy():
    """Part of the CPU-baseline leg: max |FP32 - FP64 oracle| of the GPU and
    of the reference Fortran on a reduced nonlinear grid (elevation m, flux
    m^2/s); the oracle and the reference are the checkers here."""
    from gcomcot import DeviceRun
    from gcomcot import scenarios as S
    from oracle import bind
    sc = S.nonlinear_layer(nx=256, ny=192, steps=400)
    o64 = bind.Oracle(sc, 64)
    o64.step(sc.steps)
    r64 = o64.state(1)
    d = DeviceRun(sc)
    d.run(sc.steps)
    g = d.state(1)
    d.close()
    err = lambda x, k: float(np.max(np.abs(np.asarray(x[k], np.float64) - r64[k])))
    res = {"grid": "256x192 nonlinear spherical, 400 steps",
           "gpu_fp32_vs_fp64": {k: err(g, k) for k in ("z1", "m1", "n1")}}
    if bind.Reference.available():
        ref = bind.Reference(sc)
        ref.step(sc.steps)
        rs = ref.state(1)
        res["reference_fp32_vs_fp64"] = {k: err(rs, k) for k in ("z1", "m1", "n1")}
        res["gpu_vs_reference_max_abs"] = {k: float(np.max(np.abs(
            np.asarray(g[k], np.float64) - rs[k]))) for k in ("z1", "m1", "n1")}
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: one GPU only (got --gpus {a.gpus}); scenarios would shard across GPUs "
              "whole, one process per GPU, but that is not built", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        print(json.dumps({"copy_probe": copy_probe(a.copy_probe)}))
        return
    if a.linear_probe:
        print(json.dumps({"linear_probe": linear_probe(a)}))
        return
    from gcomcot import nonlinear_layer

    sc, cells, gcups, wall, kern_ms, achieved, traffic, arrived, ms_total, tile = bench_c1(a)
    cpu = None
    if a.cpu_steps > 0:
        log("cpu baseline")
        csc = nonlinear_layer(nx=a.nx, ny=a.ny, steps=a.cpu_steps)
        cpu = cpu_baseline(csc, a.cpu_steps)
        cpu["host"] = host_cpu()
        log("cpu baseline: the restatement on all cores")
        cpu["all_cores_restatement"] = cpu_all_cores(csc, a.cpu_steps)
        if a.accuracy:
            cpu["accuracy"] = cpu_accuracy()
    others = bench_other_configs(a) if a.others else None
    outputs = bench_outputs(a) if a.outputs else None
    source = bench_source(a) if a.source else None
    batch = bench_batch(a) if a.batch else None
    serve, kept = bench_serve(a) if a.serve else (None, [])
    if cpu is not None and kept:
        log(f"cpu baseline: FP64 oracle on {len(kept)} served cases")
        cpu["served_fp64"] = cpu_served_accuracy(kept)
    copy = copy_probe(1024)

    line = {
        "metric": METRIC,
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
                   "cells": cells, "tile": list(tile), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_per_cell_update": round(traffic / cells, 2) if traffic else None,
                     "practical_peak": copy["read_plus_write_GBs"],
                     "frac_of_practical": round(achieved / copy["read_plus_write_GBs"], 4),
                     "limit": "VALU issue: ~560 VALU instructions per cell, whose stream at "
                              "one wave64 instruction per SIMD every 4 cycles would take ~179 us "
                              "per launch; the launch runs close to that, not at HBM bandwidth"},
        "cpu_baseline": cpu,
        "kernel_ms_mean": round(kern_ms, 5),
        "event_ms_total": round(ms_total, 3),
        "cells_arrived": arrived,
        "hbm_copy": copy,
        "other_configs": others,
        "outputs": outputs,
        "source": source,
        "batch": batch,
        "serve": serve,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
