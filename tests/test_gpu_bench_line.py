"""bench.py's JSON line on a small scene: one line on stdout with the fields
the measurement contract names (metric, value, unit, n_gpus, steps, warmup,
ms_per_step, higher_is_better, scaling, vs_baseline, dtype, data, config), a
roofline object (bound, achieved, peak, unit, frac, traffic) whose fraction is
achieved over peak, and the host-CPU baseline object (value, unit, cores,
kind, sample). The committed rocprofv3 counters belong to the 128^3 scene, so
at this size the line carries none of them."""

import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.gpu


def test_bench_line_contract():
    cmd = [sys.executable, "bench.py", "--side", "32", "--steps", "3", "--warmup", "12", "--early-window", "2", "5",
           "--no-batch", "--no-serve", "--no-grid", "--no-config1", "--cpu-steps", "1"]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [line for line in r.stdout.splitlines() if line.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 12
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["dtype"] == "f32" and "workload" in d["config"]
    rf = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rf, k
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0.0 < rf["frac"] < 1.0 and rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], abs=1e-5)  # 5 decimals
    assert rf["traffic"] is None and rf["valu_frac"] is None and rf["ta_busy_frac"] is None
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("port", "reference")
    assert cb["unit"] == d["unit"] or cb["unit"] == "particle-steps/s"
