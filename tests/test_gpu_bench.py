"""GPU: bench.py prints the one JSON line its contract asks for.

A small grid (512 x 384, a few steps) with the CPU-baseline leg on (one
step of the reference, or of the restatement where the reference is not
built), so every object of the line is produced; the values are checked
for consistency, not for speed.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_line_carries_the_contract():
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--nx", "512", "--ny", "384",
           "--steps", "6", "--warmup", "2", "--cpu-steps", "1", "--accuracy", "0",
           "--others", "0", "--batch", "0", "--serve", "0", "--outputs", "0", "--source", "0"]
    print("[gpu] bench.py on 512x384", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 6 and d["warmup"] == 2
    assert d["unit"] == "GCUPS" and d["dtype"] == "f32" and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0
    assert "workload" in d["config"] and d["config"]["cells"] == 512 * 384
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] in ("hbm", "mfma") and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    # 36 algorithmic bytes per cell update over the mean kernel time
    assert abs(rf["achieved"] - 36 * 512 * 384 / (d["kernel_ms_mean"] * 1e-3) / 1e9) \
        < 1e-3 * rf["achieved"] + 0.1
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["kind"] in ("reference", "port") and cb["cores"] == 1 and cb["value"] > 0
    assert cb["all_cores_restatement"]["cores"] >= 1
