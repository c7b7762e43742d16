"""Host-side checks that need no GPU: the C ABI library loads and exports what
include/fluidsx.h declares, the ctypes mirrors match the C layouts, and the
set-up entry points validate their input."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fluids3_amd import abi, derive_params, fill_points, scenes
from fluids3_amd.particles import FluidsError, SIGNATURES, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fluidsx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), f"libfluidsx.so does not export {n}"
    assert set(names) == set(SIGNATURES), "the ctypes table and the header disagree"


def test_code_object_exports_the_module_contract():
    """The gfx950 code object carries the reference's kernel names and globals
    (particles.cuh:29-42, particles.cu:17-20): their names are in the
    code object's string table."""
    co = os.path.join(ROOT, "fluids3_amd", "particles_gfx950.hsaco")
    data = open(co, "rb").read()
    assert b"gfx950" in data
    for k in ("insertParticles", "countingSortFull", "computePressure", "computeForce", "advanceParticles",
              "prefixSum", "prefixFixup", "countingSortIndex", "FParams", "FPnts", "FPntTmp", "FAccel",
              "countingSortFullRanked", "FHostCaps", "FPpt"):
        assert b"\x00" + k.encode() + b"\x00" in data, k


def test_buffer_ids_match_the_header():
    """Buffer ids and bitmask size in fluids3_amd.abi equal the #defines of
    include/fluidsx_params.h."""
    text = open(os.path.join(ROOT, "include", "fluidsx_params.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define (\w+) (\d+)\b", text)}
    names = ["FPOS", "FVEL", "FCLR", "FVEVAL", "FFORCE", "FPRESS", "FGCELL", "FGNDX", "FID", "AGRID", "AGRIDCNT",
             "AGRIDOFF", "AAUXARRAY1", "AAUXSCAN1", "AAUXARRAY2", "AAUXSCAN2", "ANBRBITS", "AFREC", "ATAILFLAG",
             "ATAILSCAN", "ATAILAUX1", "ATAILAUXS1", "ATAILAUX2", "ATAILAUXS2"]
    for k in names:
        assert getattr(abi, k) == defs[k], k
    assert abi.NBR_WORDS == defs["FX_NBR_WORDS"]
    assert abi.GRID_UNDEF == defs["FX_GRID_UNDEF"]


def test_params_layout_688_bytes():
    assert C.sizeof(abi.FParams) == 688
    assert abi.FParams.gridAdj.offset == 240 and abi.FParams.pmem.offset == 684
    P = derive_params(scenes.dambreak_16k())
    assert P.gridRes.tolist() == [29, 29, 29] and P.pnum == 16384


def test_derive_rejects_bad_scenes():
    sc = scenes.dambreak_16k()
    sc.num_points = 0
    with pytest.raises(FluidsError):
        derive_params(sc)
    sc = scenes.dambreak_16k()
    sc.grid_density = 0.3  # cell of 1.5h/0.3 = 5h: fine; 30 makes the search exceed gridAdj
    derive_params(sc)
    sc.grid_density = 30.0
    with pytest.raises(FluidsError, match="exceeds 4"):
        derive_params(sc)


def test_fill_is_deterministic_and_in_block():
    sc = scenes.dambreak_4m()
    sc.num_points = 200_000
    pos, clr = fill_points(sc)
    assert pos.shape == (200_000, 3)
    s = 0.1 * scenes.rest_spacing()
    lo, hi = np.array(sc.init_min.tolist()), np.array(sc.init_max.tolist())
    assert np.all(pos >= lo - s) and np.all(pos <= hi + s)
    assert np.array_equal(pos, fill_points(sc)[0])
    assert np.all((clr >> 24) == 255)


def test_fill_rejects_bad_volumes():
    sc = scenes.dambreak_16k()
    sc.spacing = 1e-9  # 6e10 sites per layer
    with pytest.raises(FluidsError, match="too large"):
        fill_points(sc, 10)
    sc = scenes.dambreak_16k()
    sc.init_max.x = float("nan")
    with pytest.raises(FluidsError, match="finite"):
        fill_points(sc, 10)


def test_batch_scene_counts():
    sc = scenes.batch_scenes()
    n = [s.num_points for s in sc]
    assert len(n) == 64 and min(n) >= 8192 and max(n) <= 262144
    assert len(set(round(s.visc, 6) for s in sc)) == 64
    for s in sc[:4]:
        P = derive_params(s)
        assert P.gridSrch == 3


def test_create_without_code_object_fails_cleanly():
    from fluids3_amd import Particles

    with pytest.raises(FluidsError):
        Particles(code_object="/nonexistent/particles.hsaco")


def test_bench_refuses_more_than_one_gpu():
    """Every box has one GPU: bench.py --gpus N > 1 exits with a message."""
    import subprocess
    import sys

    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode != 0
    assert "one GPU" in out.stdout + out.stderr
