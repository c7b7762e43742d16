"""bench.py's roofline object under its cost model, on hand-made stage times:
the dominant stage is chosen by time, the pair stages are priced on the FP32
vector roof (the force stage for its pair terms only, 31 flops a pair), the
streaming stages on HBM, and the figures follow from the published MI355X
peaks (157.3 TFLOP/s FP32, 8 TB/s)."""
import importlib.util
import math
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


STAGES = {"insert": 0.04, "prefix": 0.016, "sort": 0.14, "pressure": 0.95, "force": 2.0, "advance": 0.055}
N, G, C, P = 4_000_000, 101 * 53 * 53, 1114.95, 174.29


def test_force_stage_is_priced_on_its_pair_terms():
    b = load_bench()
    r = b.roofline(dict(STAGES), N, G, C, P)
    assert r["kernel"] == "computeForce" and r["bound"] == "valu" and r["unit"] == "TFLOP/s"
    assert r["peak"] == 157.3
    flops = (31.0 * P + 2.0) * N
    assert math.isclose(r["algorithmic_flops"], flops, rel_tol=1e-12)
    # the candidate tests belong to the density stage: counting them here too
    # would credit the force stage with (8C + 31P + 2) N
    assert not math.isclose(r["algorithmic_flops"], (8.0 * C + 31.0 * P + 2.0) * N, rel_tol=1e-3)
    tflops = flops / 2.0e-3 / 1e12
    assert math.isclose(r["achieved"], tflops, abs_tol=1e-3)
    assert math.isclose(r["frac"], tflops / 157.3, abs_tol=1e-4)
    assert r["algorithmic_bytes"] == 44 * N
    # two 16-byte record loads per pair, per second of the stage
    assert math.isclose(r["pair_loads_per_s"], 2.0 * P * N / 2.0e-3, rel_tol=1e-3)


def test_streaming_stage_is_priced_on_hbm():
    b = load_bench()
    st = dict(STAGES, advance=5.0)  # advance dominant: 88 bytes and 60 flops a particle
    r = b.roofline(st, N, G, C, P)
    assert r["kernel"] == "advanceParticlesClock" and r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert r["peak"] == 8000.0
    gbs = 88.0 * N / 5.0e-3 / 1e9
    assert math.isclose(r["achieved"], gbs, abs_tol=0.1)
    assert math.isclose(r["frac"], gbs / 8000.0, abs_tol=1e-4)
    assert "pair_loads_per_s" not in r


def test_cell_tables_count_in_the_sort_bytes():
    b = load_bench()
    st = dict(STAGES, sort=9.0)
    r = b.roofline(st, N, G, C, P)
    assert r["kernel"] == "countingSortFullRanked" and r["bound"] == "hbm"
    assert r["algorithmic_bytes"] == 92 * N + 4 * (G + 1)


def test_batch_launches_name_the_batch_kernels():
    b = load_bench()
    r = b.roofline(dict(STAGES), N, G, C, P, batch=True)
    assert r["kernel"] == "computeForceBatch"
    t = b.stage_table(dict(STAGES), N, G, C, P)
    assert set(t) == set(STAGES)
    # density: 20 bytes and 8C + 5P + 2 flops a particle
    assert math.isclose(t["pressure"]["TFLOP_s"], (8 * C + 5 * P + 2) * N / 0.95e-3 / 1e12, abs_tol=1e-3)
    assert math.isclose(t["pressure"]["GB_s"], 20 * N / 0.95e-3 / 1e9, abs_tol=0.1)
