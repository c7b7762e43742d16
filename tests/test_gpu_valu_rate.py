"""The VALU issue-rate probe (tests/valu_rate): FP32 add and FMA chains at
eight waves per SIMD.  Its rates are the ceilings the pair kernels' counters
are read against: an instruction of two source operands issues faster than a
three-operand FMA, and neither is above one per SIMD per clock."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "valu_rate", "valu_rate")


def test_valu_issue_rates():
    if not os.path.exists(PROBE):
        raise RuntimeError("tests/valu_rate/valu_rate is not built (run __graft_entry__.build())")
    r = subprocess.run([PROBE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    line = json.loads(r.stdout.strip().splitlines()[-1])
    print(line)
    add, fma = line["add_f32_per_simd_s"], line["fma_f32_per_simd_s"]
    # MI355X clocks at most 2.4 GHz: at most one wave64 instruction per SIMD per clock
    assert 1e8 < fma < 2.4e9 and 1e8 < add < 2.4e9
    assert add > fma
