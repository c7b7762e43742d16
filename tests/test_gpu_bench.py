"""bench.py's output: one JSON line with the benchmark fields (steps and
warm-up as asked, the metric, a roofline object for the dominant kernel and
the CPU baseline beside it), checked on the 16,384-particle dam break with a
handful of steps.  bench.py runs as a child process, as it is run from the
command line."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [ln for ln in out.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, out.stdout[-2000:]
    return json.loads(lines[0])


def test_bench_line_contract():
    d = run_bench("--workload", "dambreak_16k", "--steps", "7", "--warmup", "3", "--no-autotune")
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["steps"] == 7 and d["warmup"] == 3 and d["n_gpus"] == 1
    assert d["value"] > 0 and d["higher_is_better"] is True
    assert abs(d["value"] - 16384 / (d["ms_per_step"] * 1e-3)) <= 1e-3 * d["value"]
    assert d["config"]["particles"] == 16384 and d["config"]["workload"].startswith("dambreak_16k")
    r = d["roofline"]
    assert r["bound"] in ("hbm", "valu") and r["peak"] > 0 and 0 < r["frac"] < 1
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    c = d["cpu_baseline"]
    assert c["value"] > 0 and c["cores"] >= 1 and c["kind"] in ("port", "reference")
    assert "single_thread" in c and c["single_thread"]["cores"] == 1 and c["single_thread"]["value"] > 0
    assert c["dambreak_16k_single_thread"]["value"] > 0
    w = d["config"]["timed_window"]
    assert (w["steps"], w["warmup"], w["autotuned"]) == (7, 3, False)
