"""The two batched configurations at their full size, against the oracle.

batch64: 64 independent dam breaks, counts log-uniform in [8,192, 262,144],
viscosity uniform in [0.2, 0.5], each box filled to 0.4 (5.3M particles), run
for 100 steps as the captured step graph.  The four smallest scenes' slices
are bit-identical to the FP32 oracle run on each scene alone; all 64 scenes
pass the properties any correct step keeps: per-scene particle counts, ids a
permutation of the scene's slots, offsets the exclusive scan of the counts,
cells in sorted order, and the cell of every particle recomputed from its
position (after one more insert, scan and sort).

adjoint_batch32: 32 scenes, counts log-uniform in [16,384, 1,048,576] (9.4M
particles), each fitting its viscosity to a 40-step rollout at viscosity 0.35,
forward+adjoint with checkpoints every 8 steps in one launch per stage.  The
two smallest scenes' gradients are bit-identical to the FP32 oracle adjoint
(loss and scalar gradients to 1e-9: only the order of the FP64 sums differs),
with the FP64 adjoint's norm-wise error bounded by 1e-3 and printed; one scene
of more than 512k particles is bit-identical to the same scene's single-scene
GPU adjoint.
"""
import os

import numpy as np
import pytest

from fluids3_amd import Particles, abi, derive_params, fill_points, scenes
from fluids3_amd.rollout import target_positions
from oracle import oracle
from parity import assert_int_equal, bits_equal, rel_err

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
F32 = np.float32


def cells_from_pos(P, pos):
    """Cell of each position by the reference formula (particles.cu:34-44) in
    numpy float32, GRID_UNDEF outside [1, gridScanMax]."""
    g = (pos.astype(F32) - np.array(P.gridMin.tolist(), F32)) * np.array(P.gridDelta.tolist(), F32)
    sm = np.array(P.gridScanMax.tolist(), F32) + F32(1)
    inside = np.all((g >= F32(1)) & (g < sm), axis=1)
    c = g.astype(np.int64)  # truncation; only used where inside
    idx = (c[:, 1] * P.gridRes.z + c[:, 2]) * P.gridRes.x + c[:, 0]
    return np.where(inside, idx, abi.GRID_UNDEF)


def test_batch64_full_size_100_graph_steps():
    sc = scenes.batch_scenes(64)
    T = 100
    with Particles() as ps:
        ps.RestartBatch(sc)
        N = ps.NumPoints()
        print(f"batch64: {len(sc)} scenes, {N:,} particles, {T} graph steps", flush=True)
        ps.CaptureGraph()
        ps.LaunchGraph(T // 2)
        g = ps.state()
        ranges = [ps.SceneRange(s) for s in range(len(sc))]
        params = [ps.SceneParams(s) for s in range(len(sc))]
        # one more insert, scan and sort: cells then belong to the current positions
        ps.InsertParticles()
        ps.PrefixScanParticles()
        ps.CountingSort()
        h = ps.state()

    # the four smallest scenes against the oracle, bit for bit
    smallest = sorted(range(len(sc)), key=lambda s: sc[s].num_points)[:4]
    for s in smallest:
        p0, n, c0, nc = ranges[s]
        P = derive_params(sc[s])
        pos, clr = fill_points(sc[s])
        r = oracle.run(P, oracle.make_state(P, pos, clr=clr), T, threads=THREADS)
        sl = slice(p0, p0 + n)
        for k in ("pos", "vel", "veval"):
            assert bits_equal(g[k][sl], r[k]), (s, k)
        valid = r["gcell"] != abi.GRID_UNDEF
        for k in ("press", "force"):
            assert bits_equal(g[k][sl][valid], r[k][valid]), (s, k)
        assert_int_equal(f"id {s}", g["id"][sl] - p0, r["id"])
        gc = g["gcell"][sl]
        assert_int_equal(f"gcell {s}", np.where(gc == abi.GRID_UNDEF, gc, gc - c0), r["gcell"])
        assert_int_equal(f"cnt {s}", g["cnt"][c0:c0 + nc], r["cnt"])
        assert_int_equal(f"off {s}", g["off"][c0:c0 + nc] - p0, r["off"])
        print(f"  scene {s}: {n} particles, visc {sc[s].visc:.3f}, bit-identical to the oracle after {T} steps")

    # every scene: properties of a correct step
    for s, (p0, n, c0, nc) in enumerate(ranges):
        P = params[s]
        sl = slice(p0, p0 + n)
        cnt = h["cnt"][c0:c0 + nc].astype(np.int64)
        off = h["off"][c0:c0 + nc].astype(np.int64)
        assert cnt.sum() == n, f"scene {s}: counts sum to {cnt.sum()}, not {n}"
        assert np.array_equal(off, p0 + np.concatenate(([0], np.cumsum(cnt)[:-1]))), f"scene {s}: offsets"
        ids = np.sort(h["id"][sl].astype(np.int64))
        assert np.array_equal(ids, np.arange(p0, p0 + n)), f"scene {s}: ids are not a permutation"
        gc = h["gcell"][sl].astype(np.int64)
        nvalid = n - int(cnt[-1])  # the last slot is the scene's out-of-range tail
        assert np.all(np.diff(gc[:nvalid]) >= 0), f"scene {s}: cells out of order"
        assert np.all(gc[nvalid:] == abi.GRID_UNDEF), f"scene {s}: tail holds in-range cells"
        local = np.where(gc == abi.GRID_UNDEF, gc, gc - c0)
        assert_int_equal(f"scene {s} cells from positions", local, cells_from_pos(P, h["pos"][sl]))


def test_adjoint_batch32_full_size():
    sc = scenes.batch32_scenes()
    T, K = 40, 8
    with Particles() as ps:
        print(f"adjoint_batch32: fitting targets for {len(sc)} scenes", flush=True)
        targets = [target_positions(ps, s, T, visc=0.35, k=s.int_stiff) for s in sc]
        ps.RestartBatch(sc)
        N = ps.NumPoints()
        print(f"adjoint_batch32: {len(sc)} scenes, {N:,} particles, {T} steps, checkpoints every {K}", flush=True)
        ps.AdjointConfigure(T, K)
        assert ps.AdjointSegment() == K
        ps.AdjointSetTarget(np.concatenate(targets))
        loss, _ = ps.AdjointForward()
        gb = ps.AdjointBackward()
        ranges = [ps.SceneRange(s) for s in range(len(sc))]

        # one scene above 512k particles against its single-scene GPU adjoint
        big = next(s for s in range(len(sc)) if sc[s].num_points >= 524288)
        p0, n = ranges[big][:2]
        print(f"  scene {big}: {n} particles alone", flush=True)
        ps.Restart(sc[big])
        ps.AdjointConfigure(T, K)
        ps.AdjointSetTarget(targets[big])
        l1, _ = ps.AdjointForward()
        g1 = ps.AdjointBackward()
        for k in ("gpos", "gvel", "gveval"):
            assert bits_equal(gb[k][p0:p0 + n], g1[k]), (big, k)
        assert loss[big] == pytest.approx(l1[0], rel=1e-12)
        assert gb["gvisc"][big] == pytest.approx(g1["gvisc"][0], rel=1e-12)
        assert gb["gk"][big] == pytest.approx(g1["gk"][0], rel=1e-12)

    # the two smallest scenes against the FP32 oracle adjoint
    for s in sorted(range(len(sc)), key=lambda s: sc[s].num_points)[:2]:
        p0, n = ranges[s][:2]
        P = derive_params(sc[s])
        pos, clr = fill_points(sc[s])
        r = oracle.grad(P, {"pos": pos, "clr": clr}, T, targets[s], fp64=False, threads=THREADS)
        for k in ("gpos", "gvel", "gveval"):
            assert bits_equal(gb[k][p0:p0 + n], r[k]), (s, k, np.abs(gb[k][p0:p0 + n] - r[k]).max())
        assert loss[s] == pytest.approx(r["loss"], rel=1e-9)
        assert gb["gvisc"][s] == pytest.approx(r["gvisc"], rel=1e-9)
        assert gb["gk"][s] == pytest.approx(r["gk"], rel=1e-9)
        # the adjoint's own error: FP64 adjoint over the same FP32 forward, norm-wise
        # (element-wise it is led by particles whose gradient is near zero)
        m = oracle.grad_mixed(P, {"pos": pos, "clr": clr}, T, targets[s], threads=THREADS)
        nw = {k: float(np.linalg.norm(gb[k][p0:p0 + n] - m[k]) / np.linalg.norm(m[k]))
              for k in ("gpos", "gvel", "gveval")}
        el = {k: rel_err(gb[k][p0:p0 + n], m[k]) for k in ("gpos", "gvel", "gveval")}
        sv = abs(gb["gvisc"][s] - m["gvisc"]) / abs(m["gvisc"])
        print(f"  scene {s}: {n} particles, bit-identical to the FP32 oracle adjoint; vs FP64 adjoint "
              f"of the same forward, norm-wise: {', '.join(f'{k} {v:.1e}' for k, v in nw.items())} "
              f"(element-wise floored max {max(el.values()):.1e}), dL/dvisc {sv:.1e}")
        assert max(nw.values()) <= 1e-3 and sv <= 1e-3
