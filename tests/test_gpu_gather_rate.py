"""The scattered-gather rate probe (tests/gather_rate): every lane reads its
own 32-byte record (two 16-byte loads) from a random place in a window of
records, three records in flight per lane.  Its rates are the ceiling the
force stage's pair gathers are read against: small windows (served by the
caches) are at least as fast as the whole buffer (served by HBM)."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "gather_rate", "gather_rate")


def test_scattered_gather_rates():
    if not os.path.exists(PROBE):
        raise RuntimeError("tests/gather_rate/gather_rate is not built (run __graft_entry__.build())")
    r = subprocess.run([PROBE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    print(line)
    r64, r256, rall = line["lane_loads_per_s_64KB"], line["lane_loads_per_s_256KB"], line["lane_loads_per_s_128MB"]
    # 16-byte loads: the whole-buffer rate is bounded by HBM (8 TB/s moves at most 5e11 such loads per
    # second when each brings a 64-byte line), the cached windows by the CUs' load paths
    assert 1e10 < rall < 2e12 and 1e10 < r64 < 1e13
    assert r64 >= 0.9 * rall and r256 >= 0.9 * rall
