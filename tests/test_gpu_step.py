"""GPU parity of the SPH step (gfx950 code object through the C ABI) against the CPU oracle.

Bars: integer outputs equal; density, force and the integrated state within
1e-5 floored relative error (tests/parity.py).  The kernels and the FP32
oracle evaluate the same IEEE expression trees in the same order, so the
tests also require bit-identical FP32 results where that contract applies.
"""
import os

import numpy as np
import pytest

from fluids3_amd import Particles, abi, scenes
from oracle import oracle
from parity import (assert_close, assert_int_equal, bits_equal, cell_members_equal, rel_err,
                    valid_mask)

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def ps():
    p = Particles()
    yield p
    p.close()


def snapshot_to_oracle(st, dtype=np.float32):
    s = {k: np.array(v, copy=True) for k, v in st.items() if k != "time"}
    s["time"] = float(st["time"])
    return oracle.convert(s, dtype)


def fresh_oracle(params, s0, dtype=np.float32):
    return oracle.make_state(params, s0["pos"], vel=s0["vel"], veval=s0["veval"], clr=s0["clr"],
                             ids=s0["id"], dtype=dtype, time=s0["time"])


def compare_state(g, ref, floats=("press", "force", "pos", "vel", "veval"), bitexact=True):
    for k in ("gcell", "gndx", "id", "clr", "cnt", "off"):
        assert_int_equal(k, g[k], ref[k])
    assert cell_members_equal(g["grid"], ref["grid"], g["off"], g["cnt"]), "cell member lists differ"
    m = valid_mask(ref)
    for k in floats:
        a, b = g[k], ref[k]
        if k in ("press", "force"):
            a, b = a[m], b[m]
        assert_close(k, a, b)
        if bitexact:
            assert bits_equal(a, b), f"{k}: not bit-identical to the FP32 oracle (max rel {rel_err(a, b):.2e})"


def test_one_step_dambreak_16k(ps):
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    ps.Run(1)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 1)
    compare_state(g, ref)
    # error against the FP64 build of the same formulas, per result: stated, and
    # bounded only loosely (force components cancel, so FP32 rounding of the
    # density difference (rho - rho0) shows up at the 1e-3 level there)
    ref64 = oracle.run(p, fresh_oracle(p, s0, np.float64), 1, fp64=True)
    m = valid_mask(ref)
    for k in ("press", "force", "pos", "vel", "veval"):
        a = g[k][m] if k in ("press", "force") else g[k]
        b = ref64[k][m] if k in ("press", "force") else ref64[k]
        a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
        floor = 1e-3 * np.sqrt(np.mean(b * b))
        e = np.abs(a - b) / np.maximum(np.abs(b), floor)
        print(f"FP32 vs FP64 {k}: max {e.max():.2e} p99 {np.quantile(e, 0.99):.2e} median {np.median(e):.2e}")
        assert e.max() < (1e-5 if k in ("press", "pos") else 1e-2), k


def test_stages_dambreak_16k(ps):
    """Each per-function stage against the oracle, fed the device state before it."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    # insert: cells and counts exact, in-cell slots a permutation of 0..cnt-1
    ps.InsertParticles()
    gi = ps.state()
    ri = oracle.run(p, fresh_oracle(p, s0), 1, oracle.STAGE_INSERT)
    assert_int_equal("gcell", gi["gcell"], ri["gcell"])
    assert_int_equal("cnt", gi["cnt"], ri["cnt"])
    cell = np.where(gi["gcell"] == abi.GRID_UNDEF, p.gridTotal, gi["gcell"])
    key = np.lexsort((gi["gndx"], cell))
    first = np.concatenate([[0], np.cumsum(ri["cnt"])[:-1]])
    want = np.arange(cell.shape[0]) - np.repeat(first, ri["cnt"].astype(np.int64))
    assert_int_equal("gndx ranks", gi["gndx"][key], want)
    # scan
    ps.PrefixScanParticles()
    gs = ps.state()
    assert_int_equal("off", gs["off"], np.concatenate([[0], np.cumsum(gs["cnt"])[:-1]]).astype(np.uint32))
    # counting sort (insert + scan + sort from the initial state)
    ps.CountingSort()
    g = ps.state()
    r = oracle.run(p, fresh_oracle(p, s0), 1, oracle.STAGE_INSERT | oracle.STAGE_SCAN | oracle.STAGE_SORT)
    for k in ("gcell", "gndx", "id", "clr", "cnt", "off"):
        assert_int_equal(k, g[k], r[k])
    for k in ("pos", "vel", "veval"):
        assert bits_equal(g[k], r[k]), k
    # pressure, force, advance, each from the device's previous state
    for stage, bit, fields in ((ps.ComputePressure, oracle.STAGE_PRESSURE, ("press",)),
                               (ps.ComputeForce, oracle.STAGE_FORCE, ("force",)),
                               (ps.Advance, oracle.STAGE_ADVANCE, ("pos", "vel", "veval"))):
        r = oracle.run(p, snapshot_to_oracle(g), 1, bit)
        stage()
        g = ps.state()
        m = valid_mask(r)
        for k in fields:
            a, b = (g[k][m], r[k][m]) if k in ("press", "force") else (g[k], r[k])
            assert_close(k, a, b)
            assert bits_equal(a, b), k


def test_caller_runs_one_stage_on_the_cpu(ps):
    """The reference's per-function CPU/GPU choice (m_bGPU, particles.cpp:35)
    composed by the caller: GPU insert .. force through the per-stage calls,
    the advance on the host (here the oracle's) over the retrieved state,
    Commit of pos/vel/veval and the clock moved by dt through UpdateParams.
    Three such steps equal three all-GPU steps bit for bit."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    ps.Run(3)
    want = ps.state()
    ps.Restart(sc)
    for _ in range(3):
        for stage in (ps.InsertParticles, ps.PrefixScanParticles, ps.CountingSort, ps.ComputePressure,
                      ps.ComputeForce):
            stage()
        p = ps.params
        r = oracle.run(p, snapshot_to_oracle(ps.state()), 1, oracle.STAGE_ADVANCE)
        for buf, k in ((abi.FPOS, "pos"), (abi.FVEL, "vel"), (abi.FVEVAL, "veval")):
            ps.Commit(buf, r[k])
        p.time = np.float32(np.float32(p.time) + np.float32(p.dt))
        ps.UpdateParams(p)
    got = ps.state()
    assert got["time"] == want["time"]
    compare_state(got, want)


def test_stages_out_of_order_are_refused(ps):
    """A stage whose input tables the step has not made yet (offsets before the
    scan, the sorted cells before the sort, the neighbour bitmask before the
    density) is refused with an error instead of indexing stale tables; a graph
    that was captured but not launched ran nothing."""
    from fluids3_amd.particles import FluidsError

    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    for stage in (ps.PrefixScanParticles, ps.CountingSort, ps.ComputePressure, ps.ComputeForce, ps.CountPairs,
                  lambda: ps.SampleParticles((4, 4, 4), (-30.0, 0.0, -30.0), (30.0, 60.0, 30.0))):
        with pytest.raises(FluidsError):
            stage()
    ps.CaptureGraph()
    with pytest.raises(FluidsError):
        ps.ComputePressure()
    ps.InsertParticles()
    with pytest.raises(FluidsError):
        ps.CountingSort()  # the scan has not run
    ps.PrefixScanParticles()
    with pytest.raises(FluidsError):
        ps.ComputeForce()  # nor the sort and the density
    ps.CountingSort()
    with pytest.raises(FluidsError):
        ps.CountingSort()  # sorting the sorted set again
    ps.ComputePressure()
    ps.ComputeForce()
    ps.Advance()
    ps.ComputeForce()  # the tables still describe the current order
    ps.LaunchGraph(1)
    ps.ComputePressure()
    # the refused calls changed nothing: the run equals a plain one
    got = ps.state()
    ps.Restart(sc)
    ps.Run(3)
    want = ps.state()
    for k in ("pos", "vel", "veval", "id"):
        assert bits_equal(got[k], want[k]), k


def test_multi_step_bitexact_dambreak_16k(ps):
    """The 16,384-particle dam break for its full 100 steps, bit for bit."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    ps.Run(100)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 100, threads=THREADS)
    compare_state(g, ref)


def test_determinism(ps):
    sc = scenes.dambreak_16k()
    out = []
    for _ in range(2):
        ps.Restart(sc)
        ps.Run(5)
        out.append(ps.state())
    for k in ("pos", "vel", "veval", "press", "force"):
        assert bits_equal(out[0][k], out[1][k]), k
    for k in ("gcell", "gndx", "id", "cnt", "off"):
        assert_int_equal(k, out[0][k], out[1][k])


def test_graph_equals_eager(ps):
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    ps.Run(6)
    eager = ps.state()
    ps.Restart(sc)
    ps.CaptureGraph()
    for _ in range(3):  # replays with host work and synchronisation between them
        ps.LaunchGraph(1)
        ps.Synchronize()
        _ = ps.Retrieve(abi.FPOS)
    graph = ps.state()
    assert abs(graph["time"] - eager["time"]) == 0.0
    for k in ("pos", "vel", "veval", "press", "force"):
        assert bits_equal(eager[k], graph[k]), k
    for k in ("gcell", "gndx", "id", "cnt", "off"):
        assert_int_equal(k, eager[k], graph[k])


@pytest.mark.parametrize("threads", [64, 128, 512, 1024])
def test_block_size_invariance(ps, threads):
    """Results do not depend on the launch geometry (the force stage's word
    queue holds 16, 16, 8 or 4 LDS words per lane at 64, 128, 256 and 512
    threads per block; 1024-thread blocks evaluate pairs in place)."""
    sc = scenes.dambreak_16k()
    defaults = [ps.StageBlockSize(k) for k in range(abi.NUM_STAGES)]
    ps.Restart(sc)
    ps.Run(3)
    base = ps.state()
    ps.SetBlockSize(threads)
    try:
        ps.Restart(sc)
        ps.Run(3)
        g = ps.state()
    finally:
        for k, t in enumerate(defaults):
            ps.SetStageBlockSize(k, t)
    for k in ("pos", "vel", "veval", "press", "force"):
        assert bits_equal(base[k], g[k]), (threads, k)


@pytest.mark.parametrize("gd,srch,threads", [(1.0, 2, 64), (1.0, 2, 256), (2.5, 4, 64), (2.5, 4, 256)])
def test_stencil_widths(ps, gd, srch, threads):
    """Neighbour searches of 2 and 4 cells per axis (4 and 16 stencil rows;
    the force stage walks 2-cell rows with its word queue and 4-cell rows
    with its pair ring), bit-exact with the oracle over 3 steps."""
    sc = scenes.dambreak_16k()
    sc.grid_density = gd
    defaults = [ps.StageBlockSize(k) for k in range(abi.NUM_STAGES)]
    ps.SetBlockSize(threads)
    try:
        ps.Restart(sc)
        p = ps.params
        assert p.gridSrch == srch
        s0 = ps.state()
        ps.Run(3)
        g = ps.state()
    finally:
        for k, t in enumerate(defaults):
            ps.SetStageBlockSize(k, t)
    ref = oracle.run(p, fresh_oracle(p, s0), 3, threads=THREADS)
    compare_state(g, ref)


@pytest.mark.parametrize("threads", [64, 128, 256, 512])
def test_dense_cells_without_bitmask(ps, threads):
    """~500 particles per cell: interior particles have more candidates than
    the bitmask holds (FX_NBR_WORDS words), so the force stage tests them in
    place beside the word queue; row runs exceed one density tile, so the
    density stage takes its chunked path.  Bit-exact with the oracle over 2
    steps at every force path's block size."""
    sc = scenes.dam_break(16000, (-40, 0, -40), (40, 60, 40), (-4, 4, -4), (4, 12, 4))
    defaults = [ps.StageBlockSize(k) for k in range(abi.NUM_STAGES)]
    ps.SetBlockSize(threads)
    try:
        ps.Restart(sc)
        p = ps.params
        s0 = ps.state()
        ps.Run(2)
        g = ps.state()
    finally:
        for k, t in enumerate(defaults):
            ps.SetStageBlockSize(k, t)
    ref = oracle.run(p, fresh_oracle(p, s0), 2, threads=THREADS)
    compare_state(g, ref)
    # the first step's sorted state: some particle has more candidates than the bitmask holds
    srt = oracle.run(p, fresh_oracle(p, s0), 1, oracle.STAGE_INSERT | oracle.STAGE_SCAN | oracle.STAGE_SORT)
    assert oracle.pairs_grid(p, srt)[0].max() > 32 * abi.NBR_WORDS


def test_compressed_block_trajectory(ps):
    """64,000 particles at the 4M dam break's lattice spacing (0.7, about 45
    per cell, up to ~60 as the block collapses), 30 steps bit for bit: cells of
    more than 32 members, which the sort ranks wave-cooperatively, long density
    rows and deep force queues, over a trajectory."""
    sc = scenes.dam_break(64000, (-60, 0, -30), (60, 60, 30), (-54, 4, -14), (-26, 32, 14))
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    ps.Run(30)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 30, threads=THREADS)
    compare_state(g, ref)
    assert g["cnt"][:-1].max() > 32


@pytest.mark.parametrize("n", [1, 2, 63, 65, 1000, 4097])
def test_ragged_counts(ps, n):
    sc = scenes.dambreak_16k(n=n)
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    ps.Run(2)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 2)
    compare_state(g, ref)


def test_out_of_range_and_nan(ps):
    """GRID_UNDEF particles: counted in the tail slot, sorted after every cell in
    particle order, parked at (-1000)^3 with zero velocity."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    pos = s0["pos"].copy()
    rng = np.random.default_rng(7)
    out = rng.choice(pos.shape[0], 600, replace=False)
    pos[out[:200], 1] = -20.0                     # below the grid
    pos[out[200:400], 0] = p.gridMin.x + 0.5      # in the first cell layer (gc.x = 0)
    pos[out[400:598], 2] = 1.0e9                  # far outside
    pos[out[598:], :] = np.nan                    # not a number
    ps.Commit(abi.FPOS, pos)
    s0 = ps.state()
    ps.Run(1)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 1)
    assert g["cnt"][-1] == 600
    compare_state(g, ref)
    tail = slice(int(g["off"][-1]), None)
    assert np.all(g["pos"][tail] == -1000.0) and np.all(g["vel"][tail] == 0.0)
    assert np.all(np.diff(g["id"][tail].astype(np.int64)) > 0)


def test_large_tail_bucket_is_stable(ps):
    """More out-of-range particles than a cell's member ranking covers: the
    tail bucket is still ordered by particle index (flag scan), so runs are
    reproducible and equal to the oracle, ids included."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    pos = ps.state()["pos"].copy()
    pos[::4, 0] += 1000.0  # 4,096 particles leave the grid
    ps.Commit(abi.FPOS, pos)
    s0 = ps.state()
    ps.Run(2)
    g = ps.state()
    assert g["cnt"][-1] == 4096
    compare_state(g, oracle.run(p, fresh_oracle(p, s0), 2))
    ps.Restart(sc)
    ps.Commit(abi.FPOS, pos)
    ps.Run(2)
    again = ps.state()
    for k in ("pos", "vel", "veval", "id", "clr", "gcell"):
        assert bits_equal(g[k], again[k]), k


def test_crowded_cell_is_ordered(ps):
    """3,000 particles packed into one cell (more than the 2,048 members a cell
    of the per-lane and earlier ranking covered): the sort ranks the cell
    wave-cooperatively, by x and then particle index, and two steps equal the
    oracle bit for bit."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    pos = ps.state()["pos"].copy()
    rng = np.random.default_rng(12)
    lo = np.array([-22.3, 10.2, -9.8], np.float32)  # inside cell (5, 6, 10) of the 2.5-unit grid
    pos[:3000] = (lo + rng.uniform(0.0, 2.1, (3000, 3))).astype(np.float32)
    ps.Commit(abi.FPOS, pos)
    s0 = ps.state()
    first = oracle.run(p, fresh_oracle(p, s0), 1, oracle.STAGE_INSERT)
    assert first["cnt"][:-1].max() >= 3000  # what the first step's sort ranks
    ps.Run(2)
    g = ps.state()
    compare_state(g, oracle.run(p, fresh_oracle(p, s0), 2, threads=THREADS))


def trim_probe_scene(p, pos):
    """Probes and candidates for the row-trim test, written over the first
    particles of `pos`: each probe in a row of its own, at x fractions k/16 of
    its cell (the radius then falls on a slab boundary of the row's first and
    last cells) shifted by -3..3 ulps, its three row cells crowded (20
    members: slab starts exist), and candidates of its row at x distance
    R (1 - d), d in {0, 1e-7, 2e-7, 5e-7, 1e-6}, and the floats either side, on
    both sides.  Returns the probe indices and, per probe, its candidates."""
    F = np.float32
    rng = np.random.default_rng(5)
    g0 = np.array(p.gridMin.tolist(), np.float64)
    w = 1.0 / np.array(p.gridDelta.tolist(), np.float64)  # cell edge per axis
    R = float(np.sqrt(F(p.rd2)))  # radius in world units
    at = lambda gx, gy, gz: (g0 + np.array([gx, gy, gz]) * w).astype(F)
    k = 0
    probes, cands = [], []
    rows = [(cy, cz) for cy in range(2, 26, 3) for cz in range(2, 26, 2)]
    shots = [(f / 16.0, u) for f in (1, 5, 8, 11, 15) for u in (-3, -1, 0, 1, 3)]
    for (cy, cz), (frac, ulps) in zip(rows, shots):
        for cx in (9, 10, 11):
            for _ in range(20):
                pos[k] = at(cx + rng.uniform(0.05, 0.95), cy + rng.uniform(0.05, 0.95), cz + rng.uniform(0.05, 0.95))
                k += 1
        probe = at(10 + frac, cy + 0.5, cz + 0.5)
        probe[0] = probe[0] + F(ulps) * np.spacing(probe[0])
        pos[k] = probe
        probes.append(k)
        k += 1
        mine = []
        for d in (0.0, 1e-7, 2e-7, 5e-7, 1e-6):
            for side in (-1.0, 1.0):
                q = F(float(probe[0]) + side * R * (1.0 - d))
                for qx in (np.nextafter(q, F(-1e9)), q, np.nextafter(q, F(1e9))):
                    pos[k] = probe
                    pos[k, 0] = qx
                    mine.append(k)
                    k += 1
        cands.append(mine)
    return probes, cands


def test_row_trim_at_the_radius(ps):
    """The density stage skips the x slabs of a stencil row's first and last
    cells that lie beyond a particle's radius (a bound that allows for the
    rounding of the cell coordinate).  Probes put the radius on slab
    boundaries, to within a few ulps, with candidates of their own row at
    distances R (1 - d) and the floats either side.  Some of the candidates
    inside the radius lie in slabs that a bound without that allowance would
    skip (checked below, in float32 as the kernel computes it), so a bound that
    is too tight changes a probe's density: one step must equal the oracle
    bit for bit."""
    F = np.float32
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    p = ps.params
    pos = ps.state()["pos"].copy()
    probes, cands = trim_probe_scene(p, pos)
    # the construction's power: inside candidates in a slab that the bare bound
    # (sqrt(rd2) in grid units, no allowance) would skip
    gmin, dl, rd2 = F(p.gridMin.x), F(p.gridDelta.x), F(p.rd2)
    gxf = lambda x: (F(x) - gmin) * dl
    bare = F(np.sqrt(rd2)) * dl
    sensitive = 0
    for q, mine in zip(probes, cands):
        gp = gxf(pos[q, 0])
        cx = int(gp)
        lo = min(max(int(np.ceil(F(16) * (gp - F(cx - 1) - bare) - F(1))), 0), 16)
        hi = min(max(int(np.floor(F(16) * (gp - F(cx + 1) + bare))) + 1, 0), 16)
        for j in mine:
            dx = F(pos[q, 0] - pos[j, 0])
            if not dx * dx < rd2:
                continue
            g = gxf(pos[j, 0])
            slab = min(int((g - F(int(g))) * F(16)), 15)
            if (int(g) == cx - 1 and slab < lo) or (int(g) == cx + 1 and slab >= hi):
                sensitive += 1
    assert sensitive > 0
    ps.Commit(abi.FPOS, pos)
    s0 = ps.state()
    cells = oracle.run(p, fresh_oracle(p, s0), 1, oracle.STAGE_INSERT)
    for q in probes:  # every row cell of every probe has slab starts (more than 16 members)
        c = int(cells["gcell"][q])
        assert all(cells["cnt"][c + dx] > 16 for dx in (-1, 0, 1))
    ps.Run(1)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 1, threads=THREADS)
    compare_state(g, ref)
    print(f"row trim: {len(probes)} probes, {sensitive} inside candidates in slabs a bare bound would skip")


@pytest.mark.parametrize("n", [1, 2, 1023, 1024, 1025, 2047, 2048, 4097, 1024 * 1024 - 1, 1024 * 1024,
                               1024 * 1024 + 1, 3 * 1024 * 1024 + 7])
def test_scan_lengths(ps, n):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 16, size=n, dtype=np.uint32)
    got = ps.Scan(v)
    want = np.concatenate([[0], np.cumsum(v, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert_int_equal("scan", got, want)


def test_dambreak_4m_full_size(ps):
    """4,000,000 particles: size-independent properties of the sort, and the
    density and force stages against the oracle on the same device state."""
    sc = scenes.dambreak_4m()
    ps.Restart(sc)
    p = ps.params
    n = ps.NumPoints()
    ps.Run(2)
    ps.InsertParticles()
    ps.PrefixScanParticles()
    ps.CountingSort()
    g = ps.state()
    cnt = g["cnt"].astype(np.int64)
    assert cnt.sum() == n
    assert_int_equal("off", g["off"], np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.uint32))
    cell = np.where(g["gcell"] == abi.GRID_UNDEF, p.gridTotal, g["gcell"]).astype(np.int64)
    assert np.all(np.diff(cell) >= 0), "not sorted by cell"
    assert_int_equal("slot", g["gndx"], np.arange(n) - g["off"][cell].astype(np.int64))
    assert np.array_equal(np.sort(g["id"]), np.arange(n, dtype=np.uint32)), "ids are not a permutation"
    # cell of every sorted particle recomputed from its position (float32, particles.cu:34-44)
    pos = g["pos"]
    gmin = np.array(p.gridMin.tolist(), np.float32)
    gdel = np.array(p.gridDelta.tolist(), np.float32)
    gc = (pos - gmin) * gdel
    res = p.gridRes
    ok = (gc >= 1) & (gc < np.array(p.gridScanMax.tolist(), np.float32) + 1)
    inside = ok.all(axis=1)
    ci = gc.astype(np.int64)
    want = np.where(inside, (ci[:, 1] * res.z + ci[:, 2]) * res.x + ci[:, 0], abi.GRID_UNDEF)
    assert_int_equal("gcell from pos", g["gcell"], want)
    # density and force on the full state
    r = oracle.run(p, snapshot_to_oracle(g), 1, oracle.STAGE_PRESSURE, threads=THREADS)
    ps.ComputePressure()
    gp = ps.state()
    m = valid_mask(r)
    assert_close("press", gp["press"][m], r["press"][m])
    assert bits_equal(gp["press"][m], r["press"][m])
    r = oracle.run(p, snapshot_to_oracle(gp), 1, oracle.STAGE_FORCE, threads=THREADS)
    ps.ComputeForce()
    gf = ps.state()
    assert_close("force", gf["force"][m], r["force"][m])
    assert bits_equal(gf["force"][m], r["force"][m])
    # error against the FP64 build of the same formulas on the same sorted state,
    # per result: density, then force from the FP64 densities
    r64 = oracle.run(p, snapshot_to_oracle(g, np.float64), 1, oracle.STAGE_PRESSURE | oracle.STAGE_FORCE,
                     threads=THREADS, fp64=True)
    for k, a in (("press", gp["press"]), ("force", gf["force"])):
        a, b = np.asarray(a[m], np.float64).ravel(), np.asarray(r64[k][m], np.float64).ravel()
        floor = 1e-3 * np.sqrt(np.mean(b * b))
        e = np.abs(a - b) / np.maximum(np.abs(b), floor)
        nw = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"4M FP32 vs FP64 {k}: floored max {e.max():.2e} p99 {np.quantile(e, 0.99):.2e} "
              f"median {np.median(e):.2e} norm-wise {nw:.2e}")
        assert e.max() < (1e-5 if k == "press" else 1e-1), k


def test_dambreak_4m_benchmark_window_bitexact(ps):
    """The benchmark's own run at 4,000,000 particles, 5 eager warm-up steps
    then 20 steps as the captured step graph, against 25 oracle steps from the
    same initial state: the whole state bit for bit (cells, counts, offsets,
    membership, ids, density, force, positions and velocities)."""
    sc = scenes.dambreak_4m()
    ps.Restart(sc)
    p = ps.params
    s0 = ps.state()
    ps.Run(5)
    ps.CaptureGraph()
    ps.LaunchGraph(10)  # two steps per launch
    ps.Synchronize()
    g = ps.state()
    print("4M: 25 GPU steps done; 25 oracle steps", flush=True)
    ref = oracle.run(p, fresh_oracle(p, s0), 25, threads=THREADS)
    assert g["time"] == ref["time"]
    compare_state(g, ref)


def test_ref_beach_4m_benchmark_window_bitexact(ps):
    """The reference's own 4,000,000-particle beach scene (its published rate's
    scene: moving wall, sloped floor) as the benchmark runs it, 20 eager
    warm-up steps then steps of the captured graph (the first 80 of the
    benchmark's 200, to keep the oracle's share of the test under a minute),
    against 100 oracle steps from the same initial state: the whole state bit
    for bit.  (All 220 steps of the benchmark's run measured the same.)"""
    sc = scenes.ref_beach()
    ps.Restart(sc)
    p = ps.params
    assert ps.NumPoints() == 4_000_000 and p.bound_wall_force != 0.0
    s0 = ps.state()
    ps.Run(20)
    ps.CaptureGraph()
    ps.LaunchGraph(40)  # two steps per launch
    ps.Synchronize()
    g = ps.state()
    print("ref_beach 4M: 100 GPU steps done; 100 oracle steps", flush=True)
    ref = oracle.run(p, fresh_oracle(p, s0), 100, threads=THREADS)
    assert g["time"] == ref["time"]
    compare_state(g, ref)


def test_autotune_keeps_state_and_results(ps):
    """The tuner picks a listed block size per stage, leaves the state it
    started from, and a rollout under the tuned sizes is bit-identical to
    one under the default sizes."""
    defaults = [ps.StageBlockSize(k) for k in range(abi.NUM_STAGES)]
    sc = scenes.dambreak_16k()
    try:
        ps.Restart(sc)
        ps.Run(3)
        before = ps.state()
        res = ps.Autotune(2)
        assert all(t in (64, 128, 256, 512, 1024) for t, _ in res.values()), res
        assert all(ms > 0 for _, ms in res.values()), res
        after = ps.state()
        for k in ("pos", "vel", "veval", "id", "clr"):
            assert bits_equal(before[k], after[k]), k
        assert before["time"] == after["time"]
        assert all(ps.StagePpt(k) in (1, 2, 4) for k in range(abi.NUM_STAGES))
        ps.Run(4)
        tuned = ps.state()
        for k, t in enumerate(defaults):
            ps.SetStageBlockSize(k, t)
            ps.SetStagePpt(k, 1)
        ps.Restart(sc)
        ps.Run(7)
        plain = ps.state()
        for k in ("pos", "vel", "veval", "press", "force", "id", "gcell", "cnt", "off"):
            assert bits_equal(tuned[k], plain[k]), k
    finally:
        for k, t in enumerate(defaults):
            ps.SetStageBlockSize(k, t)
            ps.SetStagePpt(k, 1)


def test_reference_default_scene(ps):
    """The reference's default beach scene (cell 0.75 h under a 27-cell
    stencil, moving X- wall, sloped floor), 20,000 particles, 40 steps, bit
    for bit with the FP32 oracle: the wall's sine is the kernels' own
    (fluids3_amd/csrc/fx_trig.hpp), which the oracle shares."""
    sc = scenes.ref_beach(20000)
    ps.Restart(sc)
    p = ps.params
    assert p.gridSrch == 3 and p.grid_size < p.psmoothradius
    assert p.bound_wall_force != 0.0 and p.bound_slope != 0.0
    s0 = ps.state()
    ps.Run(40)
    g = ps.state()
    ref = oracle.run(p, fresh_oracle(p, s0), 40, threads=THREADS)
    compare_state(g, ref)
