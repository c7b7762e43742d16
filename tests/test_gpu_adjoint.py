"""GPU reverse mode (adjoint_gfx950 kernels through the C ABI) against the CPU
adjoint oracle and FP64 central differences.

Bars: against the FP32 oracle adjoint (same expression trees) the per-particle
gradients are bit-identical and the FP64-reduced loss and scalar gradients
agree to 1e-9 relative (only the order of the FP64 sums differs); checkpoint
spacing does not change a bit; against the FP64 adjoint of the same FP32
forward the error is at most 1e-5; against the FP64 adjoint of the FP64
rollout (which adds the FP32 trajectory's own divergence) it is printed and
bounded per result.
"""
import os

import numpy as np
import pytest

from fluids3_amd import Particles, abi, derive_params, fill_points, scenes
from fluids3_amd.rollout import target_positions
from oracle import oracle
from parity import bits_equal

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def ps():
    p = Particles()
    yield p
    p.close()


def gpu_grad(ps, scene, nsteps, every, target):
    ps.Restart(scene)
    ps.AdjointConfigure(nsteps, every)
    ps.AdjointSetTarget(target)
    loss, _ = ps.AdjointForward()
    g = ps.AdjointBackward()
    g["loss"] = loss
    return g


@pytest.mark.parametrize("n,T,every", [(3000, 12, 5), (80000, 4, 3)])
def test_adjoint_bitexact_against_fp32_oracle(ps, n, T, every):
    """Segments of 5, 5, 2 steps at 3,000 particles; at 80,000 particles the
    one-wave blocks (1,250) reach past a full round of the XCD-aware block
    order (8 x 128 blocks), so both the remapped blocks and the tail run."""
    sc = scenes.scaled_dam_break(n)
    target = target_positions(ps, sc, T)
    g = gpu_grad(ps, sc, T, every, target)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    r = oracle.grad(P, {"pos": pos, "clr": clr}, T, target, fp64=False, threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        assert bits_equal(g[k], r[k]), (k, np.abs(g[k] - r[k]).max())
    assert g["loss"][0] == pytest.approx(r["loss"], rel=1e-9)
    assert g["gvisc"][0] == pytest.approx(r["gvisc"], rel=1e-9)
    assert g["gk"][0] == pytest.approx(r["gk"], rel=1e-9)


@pytest.mark.parametrize("gd,srch", [(1.0, 2), (2.5, 4)])
def test_adjoint_stencil_widths(ps, gd, srch):
    """Searches of 2 and 4 cells per axis (the pair kernels' word queue and
    their pair ring), bit-exact with the FP32 oracle adjoint."""
    sc = scenes.scaled_dam_break(3000)
    sc.grid_density = gd
    T = 4
    P = derive_params(sc)
    assert P.gridSrch == srch
    target = target_positions(ps, sc, T)
    g = gpu_grad(ps, sc, T, 3, target)
    pos, clr = fill_points(sc)
    r = oracle.grad(P, {"pos": pos, "clr": clr}, T, target, fp64=False, threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        assert bits_equal(g[k], r[k]), (k, np.abs(g[k] - r[k]).max())
    assert g["gvisc"][0] == pytest.approx(r["gvisc"], rel=1e-9)
    assert g["gk"][0] == pytest.approx(r["gk"], rel=1e-9)


def test_adjoint_dense_block_without_bitmask(ps):
    # ~500 particles per cell: interior particles need more than FX_NBR_WORDS
    # bitmask words, so their pairs are found by testing candidates in place
    sc = scenes.dam_break(16000, (-40, 0, -40), (40, 60, 40), (-4, 4, -4), (4, 12, 4))
    T = 3
    target = target_positions(ps, sc, T)
    g = gpu_grad(ps, sc, T, 2, target)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    srt = oracle.run(P, oracle.make_state(P, pos, clr=clr), 1,
                     oracle.STAGE_INSERT | oracle.STAGE_SCAN | oracle.STAGE_SORT)
    assert oracle.pairs_grid(P, srt)[0].max() > 32 * abi.NBR_WORDS
    r = oracle.grad(P, {"pos": pos, "clr": clr}, T, target, fp64=False, threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        assert bits_equal(g[k], r[k]), (k, np.abs(g[k] - r[k]).max())
    assert g["gvisc"][0] == pytest.approx(r["gvisc"], rel=1e-9)
    assert g["gk"][0] == pytest.approx(r["gk"], rel=1e-9)


def test_adjoint_walls_slope_friction_moving_wall(ps):
    """Every wall branch (sloped floor with friction, moving X- wall, X+/Z
    walls, ceiling), initial velocities nonzero: GPU against the FP32 oracle,
    bit for bit (the moving wall's sine and cosine are the kernels' own,
    fluids3_amd/csrc/fx_trig.hpp, which the oracle shares)."""
    sc = scenes.dam_break(600, (-12, 0, -12), (12, 14, 12), (-11.5, 0.5, -11.5), (11.5, 13.5, 11.5))
    sc.bound_slope = 0.12
    sc.bound_friction = 0.3
    sc.bound_wall_force = 3.0
    sc.bound_wall_freq = 2.0
    T = 12
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    n = pos.shape[0]
    rng = np.random.default_rng(9)
    v0 = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    tgt = (pos + rng.uniform(-1, 1, (n, 3))).astype(np.float32)
    ps.Restart(sc)
    ps.Commit(abi.FVEL, v0)
    ps.Commit(abi.FVEVAL, v0)
    ps.AdjointConfigure(T, 5)
    ps.AdjointSetTarget(tgt)
    loss, _ = ps.AdjointForward()
    g = ps.AdjointBackward()
    r = oracle.grad(P, {"pos": pos, "clr": clr, "vel": v0, "veval": v0}, T, tgt, fp64=False, threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        e = np.linalg.norm(g[k] - r[k]) / np.linalg.norm(r[k])
        assert e <= 1e-4, (k, e)
        assert np.array_equal(g[k].view(np.uint32), r[k].view(np.uint32)), (k, e)
    assert loss[0] == pytest.approx(r["loss"], rel=1e-9)
    assert g["gvisc"][0] == pytest.approx(r["gvisc"], rel=1e-9)
    assert g["gk"][0] == pytest.approx(r["gk"], rel=1e-9)


def test_adjoint_checkpoint_spacing_invariance(ps):
    sc = scenes.scaled_dam_break(5000)
    T = 9
    target = target_positions(ps, sc, T)
    ref = gpu_grad(ps, sc, T, T, target)
    for every in (1, 2, 4, 0):  # 0: chosen from free memory (all 9 steps fit one segment)
        g = gpu_grad(ps, sc, T, every, target)
        assert ps.AdjointSegment() == (every if every > 0 else T)
        for k in ("gpos", "gvel", "gveval"):
            assert bits_equal(g[k], ref[k]), (every, k)
        for k in ("loss", "gvisc", "gk"):
            assert g[k][0] == ref[k][0], (every, k)


def test_adjoint_one_tape_buffer(ps, monkeypatch):
    """With one tape buffer (kept when a second does not fit the memory
    budget) the backward recomputes each segment after the adjoint of the
    next one instead of beside it: same bits."""
    sc = scenes.scaled_dam_break(5000)
    T = 9
    target = target_positions(ps, sc, T)
    ref = gpu_grad(ps, sc, T, T, target)
    monkeypatch.setenv("FX_ADJOINT_TAPE_BUFFERS", "1")
    for every in (2, 4):
        g = gpu_grad(ps, sc, T, every, target)
        for k in ("gpos", "gvel", "gveval"):
            assert bits_equal(g[k], ref[k]), (every, k)
        for k in ("loss", "gvisc", "gk"):
            assert g[k][0] == ref[k][0], (every, k)


def test_adjoint_tape_bitmask_size(ps, monkeypatch):
    """The tape's bitmask holds the words the densest stencil needs (twice
    that, at least 32); with only 4 words a slot most particles record no
    bits and both the forward force and the adjoint test their candidates in
    place: same bits, over segments recomputed from checkpoints."""
    sc = scenes.scaled_dam_break(5000)
    T = 9
    target = target_positions(ps, sc, T)
    ref = gpu_grad(ps, sc, T, 4, target)
    full = ps.AdjointMemory()
    monkeypatch.setenv("FX_ADJOINT_TAPE_WORDS", "4")
    g = gpu_grad(ps, sc, T, 4, target)
    assert ps.AdjointMemory() < full
    for k in ("gpos", "gvel", "gveval"):
        assert bits_equal(g[k], ref[k]), k
    for k in ("loss", "gvisc", "gk"):
        assert g[k][0] == ref[k][0], k
    # the system goes on from the rollout's end state as a plain run does
    ps.Run(2)
    after = ps.state()
    with Particles() as other:
        other.Restart(sc)
        other.Run(T + 2)
        want = other.state()
    for k in ("pos", "vel", "veval", "id"):
        assert bits_equal(after[k], want[k]), k


def test_adjoint_leaves_the_rollout_end_state(ps):
    sc = scenes.scaled_dam_break(4000)
    T = 7
    target = np.zeros((sc.num_points, 3), np.float32)
    gpu_grad(ps, sc, T, 3, target)
    after = ps.state()
    ps.Restart(sc)
    ps.Run(T)
    plain = ps.state()
    for k in ("pos", "vel", "veval", "id"):
        assert bits_equal(after[k], plain[k]), k
    assert after["time"] == plain["time"]


def test_adjoint_against_fp64_cpu_adjoints_16k(ps):
    """16,384 particles, 20 steps.  Two FP64 references:
    - the FP64 adjoint over the same FP32 forward (its tape promoted): isolates
      the adjoint's own FP32 rounding; bar 1e-5;
    - the FP64 adjoint of the FP64 rollout: adds the divergence of the FP32
      trajectory from the FP64 one (a few particles whose wall or neighbour
      branches switch between the two carry most of it); stated, with bars of
      1e-3 for the scalar gradients and 2e-2 norm-wise for the per-particle ones."""
    sc = scenes.scaled_dam_break(16384)
    T = 20
    target = target_positions(ps, sc, T)
    g = gpu_grad(ps, sc, T, 10, target)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    mx = oracle.grad_mixed(P, {"pos": pos, "clr": clr}, T, target, threads=THREADS)
    r = oracle.grad(P, {"pos": pos.astype(np.float64), "clr": clr}, T, target.astype(np.float64), fp64=True,
                    coef=(float(np.float32(sc.visc)), float(np.float32(sc.int_stiff))), threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        e_adj = np.linalg.norm(g[k] - mx[k]) / np.linalg.norm(mx[k])
        e_all = np.linalg.norm(g[k] - r[k]) / np.linalg.norm(r[k])
        per = np.linalg.norm(g[k] - r[k], axis=1) / np.maximum(np.linalg.norm(r[k], axis=1),
                                                               1e-3 * np.linalg.norm(r[k], axis=1).max())
        print(f"{k}: vs FP64 adjoint of the same forward {e_adj:.2e}; vs FP64 rollout {e_all:.2e} norm-wise, "
              f"per particle median {np.median(per):.2e} p99 {np.quantile(per, 0.99):.2e}")
        assert e_adj <= 1e-5, k
        assert e_all <= 2e-2, k
    for k in ("gvisc", "gk"):
        e_adj = abs(g[k][0] - mx[k]) / abs(mx[k])
        e_all = abs(g[k][0] - r[k]) / abs(r[k])
        print(f"{k}: {g[k][0]:.8e}; FP64 adjoint of the same forward {mx[k]:.8e} ({e_adj:.2e}); "
              f"FP64 rollout {r[k]:.8e} ({e_all:.2e})")
        assert e_adj <= 1e-5, k
        assert e_all <= 1e-3, k
    print(f"loss {g['loss'][0]:.10e}, FP64 rollout {r['loss']:.10e}")


@pytest.mark.parametrize("lo,hi,T", [(1500, 6000, 6), (15000, 40000, 3)])
def test_batch_adjoint_equals_single_scene_adjoints(ps, lo, hi, T):
    """The larger batch launches more than a full round of the XCD-aware block
    order (8 x 128 one-wave blocks), so logical blocks of several scenes share
    an XCD run and the per-scene FP64 partials are gathered by logical block."""
    sc = scenes.batch_scenes(count=4, lo=lo, hi=hi, seed=21)
    targets = [target_positions(ps, s, T) for s in sc]
    ps.RestartBatch(sc)
    ps.AdjointConfigure(T, 4)
    ps.AdjointSetTarget(np.concatenate(targets))
    loss, _ = ps.AdjointForward()
    gb = ps.AdjointBackward()
    ranges = []
    for s in range(len(sc)):
        p0, n, _, _ = ps.SceneRange(s)
        ranges.append((p0, n))
    single = Particles()
    try:
        for s, scene in enumerate(sc):
            p0, n = ranges[s]
            g = gpu_grad(single, scene, T, 4, targets[s])
            for k in ("gpos", "gvel", "gveval"):
                assert bits_equal(gb[k][p0:p0 + n], g[k]), (s, k)
            assert loss[s] == pytest.approx(g["loss"][0], rel=1e-12)
            assert gb["gvisc"][s] == pytest.approx(g["gvisc"][0], rel=1e-12)
            assert gb["gk"][s] == pytest.approx(g["gk"][0], rel=1e-12)
    finally:
        single.close()


def test_autograd_function_carries_the_adjoint(ps):
    import torch

    from fluids3_amd.autograd import rollout_loss

    sc = scenes.scaled_dam_break(2000)
    T = 5
    target = target_positions(ps, sc, T)
    v0 = torch.zeros((sc.num_points, 3), dtype=torch.float32, requires_grad=True)
    visc = torch.tensor(0.35, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    loss = rollout_loss(v0, visc, k, ps, sc, T, 2, target)
    (2.0 * loss).backward()
    g = gpu_grad(ps, sc, T, 2, target)
    assert float(loss.detach()) == g["loss"][0]
    expect = (g["gvel"].astype(np.float64) + g["gveval"]) * 2.0
    assert np.array_equal(v0.grad.numpy(), expect.astype(np.float32))
    assert float(visc.grad) == pytest.approx(2.0 * g["gvisc"][0], rel=1e-15)
    assert float(k.grad) == pytest.approx(2.0 * g["gk"][0], rel=1e-15)


@pytest.mark.slow
def test_adjoint_dambreak_262k_full_size(ps):
    """262,144 particles, 50 steps, checkpoints every 10.
    - The FP32 GPU gradient against the FP64 adjoint over the same FP32 forward
      (tape promoted to double), every component: bar 1e-5.
    - That FP64 adjoint against FP64 central differences of the FP64 rollout
      for 6 seeded initial-velocity components (v0 sets vel and veval), step
      1e-6 checked against 1e-6/3, and the two one-sided quotients at 1e-6/3
      checked against each other, to skip a component whose loss has a jump
      or a kink within the step (a wall band switching on, a pair crossing
      the radius, where the viscosity term is only continuous): bar 1e-4.
    - The FP32 gradient against the FP64 rollout's own adjoint: printed.  At
      this size the FP32 and FP64 trajectories part (wall-band damping switches
      on discontinuously, neighbour sets change), so the two programs'
      gradients differ by far more than rounding; no bar."""
    sc = scenes.scaled_dam_break(262144)
    T, K = 50, 10
    target = target_positions(ps, sc, T)
    g = gpu_grad(ps, sc, T, K, target)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    mx = oracle.grad_mixed(P, {"pos": pos, "clr": clr}, T, target, threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        e = np.linalg.norm(g[k] - mx[k]) / np.linalg.norm(mx[k])
        print(f"dambreak_262k {k}: FP32 GPU vs FP64 adjoint of the same forward {e:.2e}")
        assert e <= 1e-5, k
    for k in ("gvisc", "gk"):
        e = abs(g[k][0] - mx[k]) / abs(mx[k])
        print(f"dambreak_262k {k}: FP32 GPU {g[k][0]:.8e} vs FP64 adjoint of the same forward {mx[k]:.8e} ({e:.2e})")
        assert e <= 1e-5, k
    visc0, k0 = float(np.float32(sc.visc)), float(np.float32(sc.int_stiff))
    r = oracle.grad(P, {"pos": pos.astype(np.float64), "clr": clr}, T, target.astype(np.float64), fp64=True,
                    coef=(visc0, k0), threads=THREADS)
    for k in ("gpos", "gvel", "gveval"):
        print(f"dambreak_262k {k}: FP32 GPU vs FP64 rollout's adjoint {np.linalg.norm(g[k] - r[k]) / np.linalg.norm(r[k]):.2e}")
    for k in ("gvisc", "gk"):
        print(f"dambreak_262k {k}: FP32 GPU {g[k][0]:.8e}, FP64 rollout's adjoint {r[k]:.8e}")
    base = {"pos": pos.astype(np.float64), "clr": clr}
    tgt = target.astype(np.float64)
    n = sc.num_points

    def L(coef=(visc0, k0), v=None):
        s = dict(base)
        if v is not None:
            s["vel"] = v
            s["veval"] = v
        return oracle.loss(P, s, T, tgt, fp64=True, coef=coef, threads=THREADS)

    def cd(f, x0, eps):
        return (f(x0 + eps) - f(x0 - eps)) / (2 * eps)

    rv0 = r["gvel"] + r["gveval"]
    scale = np.abs(rv0).max()
    rng = np.random.default_rng(3)
    errs = []
    f0 = L(v=np.zeros((n, 3)))
    for c in rng.permutation(n * 3):
        if len(errs) == 6:
            break
        i, a = divmod(int(c), 3)

        def fv(x, i=i, a=a):
            v = np.zeros((n, 3))
            v[i, a] = x
            return L(v=v)

        e2 = 1e-6 / 3
        fp, fm = fv(e2), fv(-e2)
        d1, d2 = cd(fv, 0.0, 1e-6), (fp - fm) / (2 * e2)
        m = max(abs(d1), abs(d2), 1e-3 * scale)
        if abs(d1 - d2) > 1e-5 * m:
            continue  # a branch switch inside the step: take another component
        if abs((fp - f0) / e2 - (f0 - fm) / e2) > 1e-4 * m:
            continue  # a kink within the step: the one-sided slopes differ
        e = abs(rv0[i, a] - d2) / max(abs(d2), 1e-3 * scale)
        print(f"dambreak_262k dL/dv0[{i},{a}]: FP64 adjoint {rv0[i, a]:.8e}, central difference {d2:.8e}, rel {e:.2e}")
        errs.append(e)
    assert len(errs) == 6
    assert max(errs) <= 1e-4
