"""GPU: launch configurations do not change a bit of the results.

- particles per thread 2, 4 and 8 for the streaming stages (insert, sort,
  advance), at two block sizes, on a particle count that leaves ragged last
  blocks, against one particle per thread;
- the same with a third of the particles out of the grid (the tail bucket), for
  one scene and for a batch (one block per item and persistent);
- a batch of scenes run with persistent stage launches (a fixed grid pulling
  (scene, block) items from per-XCD queues), with one block per item, and
  with particles per thread 4, eagerly and as the captured graph.
"""
import numpy as np
import pytest

from fluids3_amd import Particles, abi, scenes

pytestmark = pytest.mark.gpu

STREAMING = (abi.STAGE_NAMES.index("insert"), abi.STAGE_NAMES.index("sort"), abi.STAGE_NAMES.index("advance"))
FIELDS = ("pos", "vel", "veval", "press", "force", "id", "gcell", "gndx", "cnt", "off", "clr")


def run_cfg(ps, restart, threads, ppt, steps, graph=False):
    for k in range(abi.NUM_STAGES):
        ps.SetStageBlockSize(k, threads)
        ps.SetStagePpt(k, ppt if k in STREAMING else 1)
    restart()
    if graph:
        ps.CaptureGraph()
        ps.LaunchGraph(steps // 2)
    else:
        ps.Run(steps)
    return ps.state()


def same(a, b, what):
    """Byte-equal fields (press and force only where the particle is in range:
    out-of-range slots are not written by the pair stages)."""
    valid = b["gcell"] != abi.GRID_UNDEF
    for k in FIELDS:
        x, y = np.ascontiguousarray(a[k]), np.ascontiguousarray(b[k])
        if k in ("press", "force"):
            x, y = x[valid], y[valid]
        assert x.dtype == y.dtype and x.shape == y.shape, (what, k)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), (what, k)


def test_particles_per_thread_is_bit_identical():
    sc = scenes.dambreak_16k(n=10007)  # ragged last blocks at every size
    with Particles() as ps:
        restart = lambda: ps.Restart(sc)
        ref = run_cfg(ps, restart, 256, 1, 5)
        for threads in (64, 256):
            for ppt in (2, 4, 8):
                print(f"ppt {ppt} at {threads} threads", flush=True)
                same(run_cfg(ps, restart, threads, ppt, 5), ref, (threads, ppt))
        with pytest.raises(Exception):
            ps.SetStagePpt(abi.STAGE_NAMES.index("force"), 2)  # pair stages take one particle per lane



def test_tail_bucket_at_every_launch_shape():
    """A third of the particles leave the grid (the insert counts the tail
    bucket once per block through LDS, blocks holding tail and in-range
    particles alike): every block size and particles-per-thread count, and the
    captured graph, give the same bits, and the bucket's count is exact."""
    sc = scenes.dambreak_16k(n=10007)
    with Particles() as ps:
        ps.Restart(sc)
        pos = ps.state()["pos"].copy()
        pos[::4, 0] += 1000.0   # 2,502 particles beyond the grid in x
        pos[1::16, 1] = -50.0   # 626 more below it
        pos[2::64] = np.nan     # 157 more that are not a number

        def restart():
            ps.Restart(sc)
            ps.Commit(abi.FPOS, pos)

        ref = run_cfg(ps, restart, 256, 1, 4)
        assert int(ref["cnt"][-1]) == 2502 + 626 + 157
        for threads, ppt, graph in ((64, 1, False), (128, 2, False), (512, 8, False), (1024, 4, False),
                                    (64, 4, True), (1024, 1, True)):
            print(f"tail bucket: {threads} threads, ppt {ppt}, graph {graph}", flush=True)
            same(run_cfg(ps, restart, threads, ppt, 4, graph), ref, (threads, ppt, graph))


def test_batch_persistent_launch_is_bit_identical():
    sc = scenes.batch_scenes(count=9, lo=500, hi=40000, seed=31)
    sc[4].num_points = 1  # a one-particle scene
    with Particles() as ps:
        restart = lambda: ps.RestartBatch(sc)
        ps.SetPersistent(False)
        ref = run_cfg(ps, restart, 256, 1, 6)
        ps.SetPersistent(True)
        for threads, ppt, graph in ((256, 1, False), (64, 4, False), (256, 4, True), (1024, 2, True)):
            print(f"persistent batch: {threads} threads, ppt {ppt}, graph {graph}", flush=True)
            same(run_cfg(ps, restart, threads, ppt, 6, graph), ref, (threads, ppt, graph))


def test_batch_tail_buckets_at_every_launch_shape():
    """Scenes of a batch with particles out of their grids (each scene's tail
    bucket counted once per block): one block per item and persistent launches,
    at several block sizes and particles per thread, eagerly and as the graph,
    give the same bits."""
    sc = scenes.batch_scenes(count=6, lo=800, hi=30000, seed=41)
    with Particles() as ps:
        ps.RestartBatch(sc)
        pos = ps.state()["pos"].copy()
        pos[::3, 0] += 1.0e4   # a third of every scene leaves its grid
        pos[1::29, 1] = np.nan

        def restart():
            ps.RestartBatch(sc)
            ps.Commit(abi.FPOS, pos)

        ps.SetPersistent(False)
        ref = run_cfg(ps, restart, 256, 1, 4)
        for persistent, threads, ppt, graph in ((False, 64, 2, False), (False, 1024, 1, True), (True, 256, 1, False),
                                                (True, 128, 4, True), (True, 1024, 2, False)):
            ps.SetPersistent(persistent)
            print(f"batch tails: persistent {persistent}, {threads} threads, ppt {ppt}, graph {graph}", flush=True)
            same(run_cfg(ps, restart, threads, ppt, 4, graph), ref, (persistent, threads, ppt, graph))

