"""GPU: a batch of independent scenes (one launch per stage over all scenes)
equals each scene run on its own, bit for bit, and scene by scene."""
import numpy as np
import pytest

from fluids3_amd import Particles, abi, scenes
from parity import assert_int_equal, bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def systems():
    a, b = Particles(), Particles()
    yield a, b
    a.close()
    b.close()


def test_batch_equals_independent_scenes(systems):
    batch, single = systems
    sc = scenes.batch_scenes(count=7, lo=600, hi=20000, seed=11)
    sc[3].num_points = 1  # a one-particle scene
    nsteps = 4
    batch.RestartBatch(sc)
    assert batch.NumScenes() == len(sc)
    batch.Run(nsteps)
    g = batch.state()
    for s, scene in enumerate(sc):
        p0, n, c0, nc = batch.SceneRange(s)
        assert n == scene.num_points
        single.Restart(scene)
        assert single.CellTableSize() == nc
        single.Run(nsteps)
        r = single.state()
        rng = slice(p0, p0 + n)
        for k in ("pos", "vel", "veval"):
            assert bits_equal(g[k][rng], r[k]), (s, k)
        valid = r["gcell"] != abi.GRID_UNDEF
        for k in ("press", "force"):
            assert bits_equal(g[k][rng][valid], r[k][valid]), (s, k)
        assert_int_equal(f"id {s}", g["id"][rng] - p0, r["id"])
        gcell = g["gcell"][rng]
        assert_int_equal(f"gcell {s}", np.where(gcell == abi.GRID_UNDEF, gcell, gcell - c0), r["gcell"])
        assert_int_equal(f"gndx {s}", g["gndx"][rng], r["gndx"])
        assert_int_equal(f"cnt {s}", g["cnt"][c0:c0 + nc], r["cnt"])
        assert_int_equal(f"off {s}", g["off"][c0:c0 + nc] - p0, r["off"])
        assert batch.SceneParams(s).pvisc == pytest.approx(scene.visc)


def test_batch_graph_equals_eager(systems):
    batch, _ = systems
    sc = scenes.batch_scenes(count=5, lo=2000, hi=30000, seed=12)
    batch.RestartBatch(sc)
    batch.Run(4)
    eager = batch.state()
    batch.RestartBatch(sc)
    batch.CaptureGraph()
    batch.LaunchGraph(2)
    graph = batch.state()
    for k in ("pos", "vel", "veval", "press", "force"):
        assert bits_equal(eager[k], graph[k]), k
