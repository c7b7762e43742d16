"""Oracle emitParticles (fluids4.0/fluid_kernels.cu:231-257 with a
counter-based draw), pinned without retyping it: the SplitMix64 hash against
its published first output, and the respawn against what the kernel must
produce whatever the draws are (distinct particles, the box's 1/256 lattice,
the fixed velocity, untouched particles unchanged)."""
import numpy as np

from oracle import oracle

BMIN = np.array([-170.0, 10.0, -20.0], np.float32)
BMAX = np.array([-190.0, 60.0, 20.0], np.float32)


def state(n, seed=3):
    rng = np.random.default_rng(seed)
    return {"pos": rng.uniform(-50, 50, (n, 3)).astype(np.float32),
            "vel": rng.uniform(-1, 1, (n, 3)).astype(np.float32),
            "veval": rng.uniform(-1, 1, (n, 3)).astype(np.float32),
            "force": rng.uniform(-1, 1, (n, 3)).astype(np.float32)}


def test_splitmix64_known_answer():
    # first output of SplitMix64 from state 0 (Steele et al. 2014; Java's SplittableRandom(0).nextLong())
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF


def respawned(s, out):
    pos, vel, veval, force = out
    return np.flatnonzero(np.any(pos != s["pos"], axis=1) | np.any(vel != s["vel"], axis=1) |
                          np.any(veval != s["veval"], axis=1) | np.any(force != s["force"], axis=1))


def test_emit_respawns_distinct_particles_on_the_box_lattice():
    s = state(1000)
    out = oracle.emit(s, 17.0, 300)
    pos, vel, veval, force = out
    idx = respawned(s, out)
    assert idx.size == 300  # distinct targets (a bijection of the thread index)
    u = (pos[idx] - BMIN) / (BMAX - BMIN) * 256.0
    assert np.all(u >= 0) and np.all(u < 256) and np.allclose(u, np.round(u), atol=1e-3)
    assert np.all(vel[idx] == np.array([5.0, -2.0, 0.0], np.float32))
    assert np.all(veval[idx] == 0) and np.all(force[idx] == 0)
    keep = np.setdiff1d(np.arange(1000), idx)
    assert np.array_equal(pos[keep], s["pos"][keep]) and np.array_equal(vel[keep], s["vel"][keep])


def test_emit_count_is_capped_and_frames_differ():
    s = state(257)
    assert respawned(s, oracle.emit(s, 1.0, 10_000)).size == 257  # every particle once
    a, b = oracle.emit(s, 1.0, 50)[0], oracle.emit(s, 2.0, 50)[0]
    assert not np.array_equal(a, b)
    assert np.array_equal(a, oracle.emit(s, 1.0, 50)[0])  # the draw depends on (frame, thread) only
