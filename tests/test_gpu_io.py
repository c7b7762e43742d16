"""GPU: .pts state files (fx_save_points / fx_load_points) in the Fluids 4.0
layout, read back independently with numpy, loaded into a fresh system, and
continued bit-identically."""
import numpy as np
import pytest

from fluids3_amd import Particles, abi, scenes
from fluids3_amd.particles import FluidsError
from parity import bits_equal

pytestmark = pytest.mark.gpu


def read_pts(path):
    """Independent reader of the layout: int32 count, int32 fields, then
    (int32 type, int32 channels, values) per field."""
    raw = open(path, "rb").read()
    n, nf = np.frombuffer(raw, np.int32, 2, 0)
    at = 8
    fields = []
    for _ in range(nf):
        ft, fc = np.frombuffer(raw, np.int32, 2, at)
        at += 8
        dt = {0: np.uint8, 1: np.int32, 2: np.float32, 3: np.float64}[int(ft)]
        cnt = int(n) * int(fc)
        fields.append((int(ft), int(fc), np.frombuffer(raw, dt, cnt, at).reshape(int(n), int(fc))))
        at += cnt * np.dtype(dt).itemsize
    assert at == len(raw)
    return int(n), fields


@pytest.fixture(scope="module")
def ps():
    p = Particles()
    yield p
    p.close()


def test_save_load_round_trip(ps, tmp_path):
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    ps.Run(3)
    st = ps.state()
    path = str(tmp_path / "state.pts")
    ps.SavePoints(path)
    n, f = read_pts(path)
    assert n == sc.num_points and [(a, b) for a, b, _ in f] == [(2, 3), (2, 3), (0, 4), (2, 3), (1, 1)]
    assert bits_equal(f[0][2], st["pos"]) and bits_equal(f[1][2], st["vel"])
    assert np.array_equal(f[2][2].reshape(-1).view(np.uint32), st["clr"])
    assert bits_equal(f[3][2], st["veval"]) and np.array_equal(f[4][2][:, 0].view(np.uint32), st["id"])
    ps.Run(2)
    cont = ps.state()
    other = Particles()
    try:
        other.Restart(sc)
        other.LoadPoints(path)
        other.Run(2)
        again = other.state()
    finally:
        other.close()
    for k in ("pos", "vel", "veval", "id", "clr", "press", "force"):
        assert bits_equal(cont[k], again[k]), k


def test_load_fluids40_file_and_count_mismatch(ps, tmp_path):
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    st = ps.state()
    n = sc.num_points
    rng = np.random.default_rng(4)
    vel = rng.uniform(-0.1, 0.1, (n, 3)).astype(np.float32)
    path = str(tmp_path / "v40.pts")
    with open(path, "wb") as fh:  # the three fields Fluids 4.0 writes
        fh.write(np.array([n, 3], np.int32).tobytes())
        fh.write(np.array([2, 3], np.int32).tobytes() + st["pos"].astype(np.float32).tobytes())
        fh.write(np.array([2, 3], np.int32).tobytes() + vel.tobytes())
        fh.write(np.array([0, 4], np.int32).tobytes() + st["clr"].astype(np.uint32).tobytes())
    ps.LoadPoints(path)
    got = ps.state()
    assert bits_equal(got["vel"], vel) and bits_equal(got["veval"], vel)
    assert np.array_equal(got["id"], np.arange(n, dtype=np.uint32))
    small = scenes.scaled_dam_break(1000)
    ps.Restart(small)
    with pytest.raises(FluidsError):
        ps.LoadPoints(path)


def test_ids_must_be_a_permutation(ps, tmp_path):
    """FID indexes the adjoint's target and gradients: an upload or a .pts
    file whose ids are not a permutation of 0..n-1 is refused, and a refused
    file leaves the state as it was."""
    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    n = sc.num_points
    ids = np.arange(n, dtype=np.uint32)
    for bad in (np.where(ids == 7, 8, ids), np.where(ids == 7, n, ids), ids[:100]):
        with pytest.raises(FluidsError):
            ps.Commit(abi.FID, bad.astype(np.uint32))
    ps.Commit(abi.FID, ids[::-1].copy())
    assert np.array_equal(ps.Retrieve(abi.FID), ids[::-1])
    ps.Restart(sc)
    before = ps.state()
    path = str(tmp_path / "dup.pts")
    ps.SavePoints(path)
    raw = bytearray(open(path, "rb").read())
    tail = len(raw) - 4 * n  # the id field is the file's last n int32
    raw[tail:tail + 4] = np.uint32(1).tobytes()  # id 0 becomes a second 1
    open(path, "wb").write(bytes(raw))
    with pytest.raises(FluidsError):
        ps.LoadPoints(path)
    after = ps.state()
    for k in ("pos", "vel", "veval", "clr", "id"):
        assert bits_equal(after[k], before[k]), k


def test_sample_particles_against_oracle(ps):
    from oracle import oracle
    from parity import rel_err

    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    ps.Run(3)
    st = ps.state()
    res, lo, hi = (24, 20, 16), (-30.0, 0.0, -30.0), (30.0, 60.0, 30.0)
    g = ps.SampleParticles(res, lo, hi, 2.0)
    r = oracle.sample(ps.params, st, res, lo, hi, 2.0)
    assert g.shape == (16, 20, 24) and np.count_nonzero(r) > 100
    assert np.array_equal(g == 0, r == 0)
    assert rel_err(g, r) <= 1e-5  # expf on the device against the host's expf


def test_sample_particles_single_particle_closed_form(ps):
    # one particle: voxels within the radius hold gausskern * exp(-r^2 / (2 r2 / 8)) * scalar
    sc = scenes.dam_break(1, (-30, 0, -30), (30, 60, 30), (0, 20, 0), (1, 21, 1))
    ps.Restart(sc)
    p = ps.params
    x0 = np.array([[0.3, 20.2, -0.1]], np.float32)
    ps.Commit(abi.FPOS, x0)
    ps.InsertParticles()
    ps.PrefixScanParticles()
    ps.CountingSort()
    res, lo, hi = (16, 16, 16), (-4.0, 16.0, -4.0), (4.0, 24.0, 4.0)
    g = ps.SampleParticles(res, lo, hi, 1.0)
    idx = np.indices((16, 16, 16)).astype(np.float64)  # z, y, x
    vx = lo[0] + idx[2] / 16 * (hi[0] - lo[0])
    vy = lo[1] + idx[1] / 16 * (hi[1] - lo[1])
    vz = lo[2] + idx[0] / 16 * (hi[2] - lo[2])
    dsq = (vx - x0[0, 0]) ** 2 + (vy - x0[0, 1]) ** 2 + (vz - x0[0, 2]) ** 2
    r2 = dsq * p.d2
    expect = np.where((dsq < p.rd2) & (dsq > 0), p.gausskern * np.exp(-r2 / (2 * p.r2 / 8)), 0.0)
    inside = expect > 0
    assert inside.sum() > 50
    assert np.allclose(g[inside], expect[inside], rtol=1e-5, atol=0)
    assert np.all(g[~inside] == 0)


@pytest.mark.parametrize("count", [1, 300, 20000])
def test_emit_particles_against_oracle(ps, count):
    """emitParticles bit-exact with the oracle's respawn (the same draws),
    then one more step runs on the respawned state."""
    from oracle import oracle

    sc = scenes.dambreak_16k()
    ps.Restart(sc)
    ps.Run(2)
    st = ps.state()
    ps.EmitParticles(7.0, count)
    g = ps.state()
    pos, vel, veval, force = oracle.emit(st, 7.0, count)
    for k, r in (("pos", pos), ("vel", vel), ("veval", veval), ("force", force)):
        assert bits_equal(g[k], r), k
    moved = np.any(g["pos"] != st["pos"], axis=1)
    assert moved.sum() == min(count, 16384)
    ps.Run(1)
    assert np.isfinite(ps.state()["pos"]).all()


def test_scene_file_scene_steps_like_the_oracle():
    """A scene read from tests/golden/scene_schema.xml through the C ABI's
    reader (sloped floor, moving X- wall, overridden viscosity) steps on the
    GPU bit-identically to the oracle for 5 steps."""
    import os

    from fluids3_amd import derive_params
    from fluids3_amd.scene_file import load_scene
    from oracle import oracle

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scene_schema.xml")
    with open(path) as fh:
        _, sc, _ = load_scene(fh.read(), 7)
    assert sc.bound_slope > 0 and sc.bound_wall_force > 0
    with Particles() as ps:
        ps.Restart(sc)
        s0 = ps.state()
        ps.Run(5)
        g = ps.state()
    P = derive_params(sc)
    ref = oracle.run(P, oracle.make_state(P, s0["pos"], clr=s0["clr"], ids=s0["id"]), 5)
    for k in ("gcell", "gndx", "id", "cnt", "off"):
        assert np.array_equal(g[k], ref[k]), k
    for k in ("pos", "vel", "veval"):
        assert bits_equal(g[k], ref[k]), k
    valid = ref["gcell"] != abi.GRID_UNDEF
    for k in ("press", "force"):
        assert bits_equal(g[k][valid], ref[k][valid]), k
