#!/usr/bin/env python3
"""Throughput of the SPH step on one MI355X: particle-steps/s at 4,000,000 particles.

Workload (default): the 4,000,000-particle dam break in the box
(-120,0,-60)..(120,120,60) with the initial block filling the left 40 % of the
box (0.40x rest spacing, about 15x rest density), FP32.  --warmup W untimed
steps, then --steps K timed steps (the configuration states 20 + 200, which
is also what runs without flags; with other values the line also carries the
20 + 200 window as "configured_window").  Inputs are resident in HBM before the timed region.

Timed region: the K steps are captured as one HIP graph, with a HIP event
before every stage and after the last, on the library's stream, and launched
once; value = N * K / elapsed.  The per-stage event pairs of that same region
give each kernel's mean duration, from which the roofline fraction of the
dominant stage is computed with the algorithmic cost model:
  bytes per particle-step: insert 20, sort 92, density 20, force 44,
      advance 88 (264 in all), plus 16 per grid cell;
  flops per particle-step: insert 6, density 8C+5P+2, force 31P+2 (the
      candidate tests are done once, by the density stage, whose bitmask the
      force stage walks), advance 60;
C (stencil candidates) and P (in-radius pairs) per particle are measured on
the same trajectory (the step is deterministic) every --stats-every steps.
The roof of the pair stages is FP32 vector ALU issue ("valu", 157.3 TFLOP/s).
Two dispatches of the no-op kernel fxMark bracket the timed region, so that a
rocprofv3 kernel trace or counter run of this command can be cut to exactly
its dispatches (--summarize).

CPU baseline: the FP32 CPU oracle (a port of the reference's step, g++ -O3
-march=x86-64-v3) on the host: all cores on the state at the start of the
timed window, one thread on one step of it, and one thread on the reference's
CPU configuration (the 16,384-particle dam break, 100 steps).

Other workloads (--workload): dambreak_4m_rest (the same 4M particles at rest
spacing), batch64 (64-scene batch, with the same scenes stepped one at a time
beside it), adjoint_262k / adjoint_batch32 (forward+adjoint rollouts), ref_beach (the
reference's default scene, the only one with a published rate:
vs_baseline = value / 6.4e8, RTX 4090, the reference's README.md:12-19).  The default line
also carries batch64, adjoint_262k and adjoint_batch32 summaries ("extras";
--no-extras leaves them out).

--summarize reduces rocprofv3 CSV output of a bench.py run to its timed
window (see summarize()).
"""
import argparse
import csv
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (and matrix) peak
NUM_CU, NUM_XCD, NUM_SIMD = 256, 8, 1024  # MI355X: per-CU counters are summed over CUs, GRBM_GUI_ACTIVE over XCDs

# Fluids 5.0 at 4,000,000 particles on its default scene: 160 fps on an RTX 4090 (the reference's README.md:12-19)
REF_BEACH_RTX4090 = 160 * 4_000_000

TRAFFIC_FILES = {"dambreak_4m": "traffic_dambreak_4M_literal.json", "dambreak_4m_rest": "traffic_dambreak_4M_rest.json",
                 "batch64": "traffic_batch64_dambreak.json", "ref_beach": "traffic_ref_beach_4M.json"}

BYTES_PER_PARTICLE = {"insert": 20, "prefix": 0, "sort": 92, "pressure": 20, "force": 44, "advance": 88}
BYTES_PER_CELL = {"insert": 4, "prefix": 8, "sort": 4, "pressure": 0, "force": 0, "advance": 0}
STAGES = ("insert", "prefix", "sort", "pressure", "force", "advance")  # the library's stage ids, in order
STAGE_KERNEL = {"insert": "insertParticles", "prefix": "prefixSum", "sort": "countingSortFullRanked",
                "pressure": "computePressure", "force": "computeForce", "advance": "advanceParticlesClock"}
# the same stages of a batch of scenes (one launch per stage over every scene)
BATCH_KERNEL = {"insert": "insertParticlesBatch", "prefix": "prefixSum", "sort": "countingSortFullBatch",
                "pressure": "computePressureBatch", "force": "computeForceBatch", "advance": "advanceParticlesBatch"}
# dambreak_4m as configured: 20 warm-up steps, 200 timed
CONFIGURED = {"dambreak_4m": (20, 200)}


def flops_per_particle(stage, C, P):
    # The candidate tests (8 flops per candidate) are done once per step, by the
    # density stage, whose in-radius bitmask the force stage walks: the force
    # stage is credited with its pair terms only (31 per pair, the pressure
    # term hoisted), so that no flop of the step is counted twice.
    return {"insert": 6.0, "prefix": 0.0, "sort": 0.0, "pressure": 8 * C + 5 * P + 2,
            "force": 31 * P + 2, "advance": 60.0}[stage]


def describe(name, sc, N, G):
    if isinstance(sc, list):
        n = [s.num_points for s in sc]
        return (f"{name}: {len(sc)} independent dam breaks in one batch, {N:,} particles "
                f"({min(n):,}..{max(n):,} per scene), {G:,} grid cells, one launch per stage")
    b0, b1 = sc.bound_min.tolist(), sc.bound_max.tolist()
    i0, i1 = sc.init_min.tolist(), sc.init_max.tolist()
    r = lambda v: "(" + ",".join(f"{x:g}" for x in v) + ")"
    kind = ("the reference's default scene (Fluids 5.0 example 2, beach front with moving wall; "
            "its published rate is measured on it)" if name == "ref_beach" else "dam break")
    return (f"{name}: {kind}, {N:,} particles, box {r(b0)}..{r(b1)}, block {r(i0)}..{r(i1)}, "
            f"{G:,} grid cells")


def stage_table(stage_ms, N, G, C, P):
    stages = {}
    for k, ms in stage_ms.items():
        b = BYTES_PER_PARTICLE[k] * N + BYTES_PER_CELL[k] * (G + 1)
        f = flops_per_particle(k, C, P) * N
        stages[k] = {"ms": round(ms, 5), "GB_s": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                     "TFLOP_s": round(f / (ms * 1e-3) / 1e12, 3) if ms > 0 else None}
    return stages


def roofline(stage_ms, N, G, C, P, batch=False):
    """Roof of the dominant stage under the cost model: FP32 vector issue or HBM."""
    dom = max(stage_ms, key=stage_ms.get)
    fl = flops_per_particle(dom, C, P) * N
    by = BYTES_PER_PARTICLE[dom] * N + BYTES_PER_CELL[dom] * (G + 1)
    t_dom = stage_ms[dom] * 1e-3
    if fl / (FP32_PEAK_TFLOPS * 1e12) >= by / (HBM_PEAK_GBS * 1e9):
        ach = fl / t_dom / 1e12
        roof = {"bound": "valu", "compute_unit": "valu-fp32", "achieved": round(ach, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4)}
    else:
        ach = by / t_dom / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4)}
    roof["kernel"] = (BATCH_KERNEL if batch else STAGE_KERNEL)[dom]
    roof["kernel_ms"] = round(stage_ms[dom], 5)
    roof["algorithmic_bytes"] = by
    roof["algorithmic_flops"] = fl
    if dom == "force":
        # its two 16-byte record loads per pair, per second: read against the
        # scattered-gather probe (tests/gather_rate), 1.03e12 from L1
        roof["pair_loads_per_s"] = float(f"{2.0 * P * N / t_dom:.4g}")
    return roof


def count_pairs_run(ps, restart, W, K, every):
    """C and P per particle on the trajectory of W warm-up + K steps (untimed, same steps)."""
    restart()
    ps.Run(W)
    csum = psum = samples = 0
    for s in range(K):
        if s % max(1, every) == 0:
            ps.InsertParticles()
            ps.PrefixScanParticles()
            ps.CountingSort()
            c, q = ps.CountPairs()
            csum += c
            psum += q
            samples += 1
            ps.ComputePressure()
            ps.ComputeForce()
            ps.Advance()
        else:
            ps.Run(1)
    ps.Synchronize()
    N = ps.NumPoints()
    return csum / (samples * N), psum / (samples * N)


def progress(msg):
    """One line per phase on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, start, N, workload, args):
    """The FP32 oracle on the host: all cores on the state at the start of the
    timed window (as many steps as fill --cpu-seconds), one thread on one step
    of it, and one thread on the reference's CPU configuration."""
    from fluids3_amd import derive_params, fill_points, scenes
    from oracle import oracle

    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    mk = lambda: oracle.make_state(p, start["pos"], vel=start["vel"], veval=start["veval"], clr=start["clr"],
                                   ids=start["id"], time=start["time"])
    st = mk()
    if args.cpu_steps > 0:
        steps = args.cpu_steps
        sec = oracle.time_steps(p, st, steps, threads=threads)
    else:  # one step, then as many more as fill the time asked for
        sec = oracle.time_steps(p, st, 1, threads=threads)
        more = max(0, min(49, int(math.ceil((args.cpu_seconds - sec) / max(sec, 1e-6)))))
        steps = 1 + more
        if more:
            sec += oracle.time_steps(p, st, more, threads=threads)
    out = {"value": round(N * steps / sec, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} consecutive FP32 oracle step(s) of the same {N:,}-particle {workload} state "
                     f"from the start of the timed window on {threads} host threads ({sec:.2f} s)",
           "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off, OpenMP (oracle/Makefile)",
           "host_cpu": host_cpu(), "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    if args.no_cpu_single:
        return out
    s1 = oracle.time_steps(p, mk(), 1, threads=1)
    out["single_thread"] = {"value": round(N / s1, 1), "cores": 1,
                            "sample": f"1 FP32 oracle step of the same state on 1 thread ({s1:.2f} s)"}
    sc0 = scenes.dambreak_16k()
    p0 = derive_params(sc0)
    pos0, clr0 = fill_points(sc0)
    s0 = oracle.time_steps(p0, oracle.make_state(p0, pos0, clr=clr0), 100, threads=1)
    out["dambreak_16k_single_thread"] = {
        "value": round(sc0.num_points * 100 / s0, 1), "cores": 1,
        "sample": f"the reference CPU path's configuration: 100 FP32 oracle steps of the 16,384-particle dam "
                  f"break from its initial lattice on 1 thread ({s0:.2f} s)"}
    return out


def traffic_for(workload, kernel, window, launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload, only if that summary was taken over the same window (timed
    steps, warm-up steps, graph or not) and the kernel ran with the same
    (threads per block, particles per thread).  Whether the block sizes came
    from the tuner or were given does not matter: the counter runs pin them."""
    tf = os.path.join(ROOT, "profiles", TRAFFIC_FILES.get(workload, "-"))
    if not os.path.exists(tf):
        return None, None
    with open(tf) as fh:
        t = json.load(fh)
    keys = ("steps", "warmup", "graph")
    tw = t.get("window") or {}
    if any(tw.get(k) != window.get(k) for k in keys) or kernel not in t.get("bytes_per_launch", {}):
        return None, None
    if list(t.get("launch", {}).get(kernel, [])) != list(launch):
        return None, None
    return round(t["bytes_per_launch"][kernel]), "profiles/" + os.path.basename(tf)


def limits_for(workload, kernel, window, launch):
    """The hardware limits of `kernel` measured by counters in the committed
    summary of this workload (texture-path busy fraction, VALU issue rate), on
    the same terms as traffic_for."""
    if traffic_for(workload, kernel, window, launch)[0] is None:
        return None
    with open(os.path.join(ROOT, "profiles", TRAFFIC_FILES[workload])) as fh:
        return json.load(fh).get("limits", {}).get(kernel)


def timed_steps(ps, K, use_graph):
    """K steps in one graph-timed region, bracketed by fxMark dispatches (tags 1 and 2)."""
    ps.Synchronize()
    ps.Mark(1)
    ms_total, stage_ms = ps.BenchSteps(K, use_graph=use_graph)
    ps.Mark(2)
    ps.Synchronize()
    return ms_total, stage_ms


def plain_bench(args, workload, K, W, extras=True, cpu=True):
    from fluids3_amd import Particles, scenes

    batch = workload == "batch64"
    ps = Particles() if args.code_object is None else Particles(args.code_object)
    ps.SetPersistent(args.persistent == "on")
    if args.stage_threads:
        for k, v in enumerate(args.stage_threads.split(",")):
            ps.SetStageBlockSize(k, int(v))
    if batch:
        # 64 independent dam breaks, counts log-uniform in [8192, 262144],
        # viscosity uniform in [0.2, 0.5], each box filled to 0.4
        sc = scenes.batch_scenes(64)
        restart = lambda: ps.RestartBatch(sc)
    else:
        sc = scenes.PRESETS[workload]()
        restart = lambda: ps.Restart(sc)
    restart()
    N, G = ps.NumPoints(), ps.CellTableSize() - ps.NumScenes()
    p = ps.params
    progress(f"{workload}: {N:,} particles, {W} warm-up + {K} timed steps")

    # warm-up, block-size tuning on the warmed-up state (the tuner restores
    # it), then the timed region (one graph of K steps with per-stage events;
    # eager launches with the same events if the graph cannot carry them)
    use_graph = not args.no_graph
    ps.Run(W)
    ps.Synchronize()
    tuned = None
    if not args.no_autotune and not args.stage_threads:
        tuned = {k: v[0] for k, v in ps.Autotune(5).items()}
    try:
        ms_total, stage_ms = timed_steps(ps, K, use_graph)
    except Exception as e:
        if not use_graph:
            raise
        print(f"graph-timed region unavailable ({e}); timing eager launches", file=sys.stderr)
        use_graph = False
        restart()
        ps.Run(W)
        ms_total, stage_ms = timed_steps(ps, K, False)
    if not use_graph:
        stage_ms = dict(stage_ms)
    thr = [ps.StageBlockSize(k) for k in range(6)]
    value = N * K / (ms_total * 1e-3)
    progress(f"{workload}: timed region {ms_total:.2f} ms; counting candidates and pairs")
    ms_step = ms_total / K

    # same trajectory again (bit-reproducible), untimed, for C and P
    want_cpu = cpu and not args.no_cpu_baseline and not batch
    if want_cpu:
        restart()
        ps.Run(W)
        start_state = ps.state()
    C, P = count_pairs_run(ps, restart, W, K, args.stats_every)

    stages = stage_table(stage_ms, N, G, C, P)
    roof = roofline(stage_ms, N, G, C, P, batch)
    window = {"steps": K, "warmup": W, "autotuned": tuned is not None, "graph": use_graph}
    si = STAGES.index(next(k for k, v in (BATCH_KERNEL if batch else STAGE_KERNEL).items() if v == roof["kernel"]))
    roof["traffic"], src = traffic_for(workload, roof["kernel"], window, (thr[si], ps.StagePpt(si)))
    if src:
        roof["traffic_source"] = src
        lim = limits_for(workload, roof["kernel"], window, (thr[si], ps.StagePpt(si)))
        if lim:
            roof["limits"] = lim

    out = {
        "metric": ("particle-steps/sec at 4M particles (1 GPU); fwd+adjoint step time vs plain"
                   if workload == "dambreak_4m"
                   else "particle-steps/sec at 4M particles (1 GPU)" if N == 4_000_000 and not batch
                   else "particle-steps/sec"),
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        # the reference's published rate (its README.md:12-19, RTX 4090) is measured on its
        # default beach scene: only that workload is compared with it
        "vs_baseline": round(value / REF_BEACH_RTX4090, 3) if workload == "ref_beach" else None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": describe(workload, sc, N, G), "particles": N, "grid_cells": G,
                   "graph": use_graph, "stage_threads": thr,
                   "stage_ppt": [ps.StagePpt(k) for k in range(6)], "persistent": args.persistent,
                   "autotuned": tuned is not None,
                   "candidates_per_particle": round(C, 2), "pairs_per_particle": round(P, 2), "stages": stages,
                   # the timed region's dispatches in a trace of this command: those between the
                   # two fxMark dispatches (K of each stage kernel)
                   "timed_window": {**window, "delimiters": "fxMark(1) .. fxMark(2)"}},
        "roofline": roof,
        "cpu_baseline": None,
    }

    # the batch's reference point: the same scenes stepped one at a time, each
    # as its own single-scene system (same warm-up, same graph-timed region)
    if batch and not args.no_single:
        one = Particles()
        tot_ms = 0.0
        for scene in sc:
            one.Restart(scene)
            one.Run(W)
            one.Synchronize()
            ms, _ = one.BenchSteps(K, use_graph=use_graph)
            tot_ms += ms
        one.close()
        v1 = N * K / (tot_ms * 1e-3)
        out["config"]["single_scene"] = {
            "value": round(v1, 1), "ms_per_step": round(tot_ms / K, 5), "batch_over_single": round(value / v1, 3),
            "how": f"the same {len(sc)} scenes run one after another, each as a single-scene system"}

    if extras and workload in CONFIGURED and (W, K) != CONFIGURED[workload]:
        cw, ck = CONFIGURED[workload]
        progress(f"{workload}: the configured window, {cw} + {ck} steps")
        restart()
        ps.Run(cw)
        cms, cstage = timed_steps(ps, ck, use_graph)
        cC, cP = count_pairs_run(ps, restart, cw, ck, 20)
        croof = roofline(cstage, N, G, cC, cP)
        out["configured_window"] = {
            "warmup": cw, "steps": ck, "value": round(N * ck / (cms * 1e-3), 1), "ms_per_step": round(cms / ck, 5),
            "candidates_per_particle": round(cC, 2), "pairs_per_particle": round(cP, 2),
            "stages": stage_table(cstage, N, G, cC, cP),
            "roofline": {k: croof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "kernel", "kernel_ms")},
            "how": "the same system and block sizes, restarted: the configuration's own warm-up and timed steps"}

    # forward+adjoint rollout beside the plain step, same start state and window
    if workload == "dambreak_4m" and args.adjoint_steps > 0:
        progress(f"{workload}: forward+adjoint of {args.adjoint_steps} steps")
        seg = args.checkpoint_every if args.checkpoint_every is not None else 0
        out["adjoint"] = adjoint_ratio(ps, restart, W, args.adjoint_steps, seg, use_graph)

    if want_cpu:
        progress(f"{workload}: CPU baseline")
        out["cpu_baseline"] = cpu_baseline(p, start_state, N, workload, args)
    ps.close()
    return out


def adjoint_ratio(ps, restart, W, T, K, use_graph):
    """GPU time of a T-step forward+adjoint rollout (checkpoints every K) over
    the time of the same T plain steps, both from the state after W warm-up
    steps.  The target is the start state's positions (the cost does not
    depend on it)."""
    from fluids3_amd import abi

    restart()
    ps.Run(W)
    ps.Synchronize()
    plain_ms, _ = ps.BenchSteps(T, use_graph=use_graph)
    restart()
    ps.Run(W)
    ps.Synchronize()
    ids = ps.Retrieve(abi.FID).astype(np.int64)
    tgt = np.empty((ps.NumPoints(), 3), np.float32)
    tgt[ids] = ps.Retrieve(abi.FPOS)
    ps.AdjointConfigure(T, K)
    K = ps.AdjointSegment()
    ps.AdjointSetTarget(tgt)
    _, f_ms = ps.AdjointForward()
    g = ps.AdjointBackward()
    tot = f_ms + g["ms"]
    return {"steps": T, "checkpoint_every": K, "plain_ms": round(plain_ms, 4), "forward_ms": round(f_ms, 4),
            "backward_ms": round(g["ms"], 4), "fwd_adjoint_over_plain": round(tot / plain_ms, 3),
            "fwd_adjoint_particle_steps_per_s": round(ps.NumPoints() * T / (tot * 1e-3), 1),
            "tape_and_checkpoint_bytes": ps.AdjointMemory()}


def adjoint_bench(args, workload, T=None, K=None, auto=True):
    """adjoint_262k: one differentiable dam break; adjoint_batch32: a batch of 32.  value =
    particle-steps/s of the forward+adjoint rollout (loss and gradients with
    respect to initial velocities, viscosity and stiffness)."""
    from fluids3_amd import Particles, scenes
    from fluids3_amd.rollout import target_positions

    ps = Particles()
    if workload == "adjoint_262k":
        sc = scenes.dambreak_262k()
        T = T if T is not None else 50
        K = K if K is not None else 10
        target = target_positions(ps, sc, T)
        restart = lambda: ps.Restart(sc)
        desc = (f"adjoint_262k: differentiable dam break, {sc.num_points:,} particles, {T} steps, checkpoints every {K}, "
                f"target = the rollout at viscosity 0.30 and stiffness 1.2")
    else:
        sc = scenes.batch32_scenes()
        T = T if T is not None else 40
        K = K if K is not None else 8
        # each scene fits its viscosity to the rollout at viscosity 0.35
        target = np.concatenate([target_positions(ps, s, T, visc=0.35, k=s.int_stiff) for s in sc])
        restart = lambda: ps.RestartBatch(sc)
        n = [s.num_points for s in sc]
        desc = (f"adjoint_batch32: {len(sc)} differentiable dam breaks in one batch, {sum(n):,} particles "
                f"({min(n):,}..{max(n):,} per scene), {T} steps, checkpoints every {K}, "
                f"target = each scene's rollout at viscosity 0.35")
    restart()
    N = ps.NumPoints()
    # plain reference: the same T steps from the same start
    ps.Run(2)
    ps.Synchronize()
    restart()
    try:
        plain_ms, _ = ps.BenchSteps(T, use_graph=not args.no_graph)
    except Exception as e:
        print(f"graph-timed region unavailable ({e}); timing eager launches", file=sys.stderr)
        restart()
        plain_ms, _ = ps.BenchSteps(T, use_graph=False)

    def rollout(k):
        restart()
        ps.AdjointConfigure(T, k)
        k = ps.AdjointSegment()
        ps.AdjointSetTarget(target)
        loss, f_ms = ps.AdjointForward()
        g = ps.AdjointBackward()
        g["bytes"] = ps.AdjointMemory()
        return k, loss, f_ms, g

    K, loss, f_ms, g = rollout(K)
    desc = desc.replace("checkpoints every 0", f"checkpoints every {K}")
    tot = f_ms + g["ms"]
    out = {
        "metric": "forward+adjoint particle-steps/sec",
        "value": round(N * T / (tot * 1e-3), 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": T,
        "warmup": 0,
        "ms_per_step": round(tot / T, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": desc, "particles": N, "checkpoint_every": K,
                   "plain_ms": round(plain_ms, 4), "forward_ms": round(f_ms, 4), "backward_ms": round(g["ms"], 4),
                   "fwd_adjoint_over_plain": round(tot / plain_ms, 3),
                   "tape_and_checkpoint_bytes": g["bytes"],
                   "loss": [round(float(x), 6) for x in loss[:8]],
                   "dL_dvisc": [float(x) for x in g["gvisc"][:8]], "dL_dstiff": [float(x) for x in g["gk"][:8]]},
        "roofline": None,
        "cpu_baseline": None,
    }
    if auto:
        # the same rollout with segments chosen from free device memory, reported beside it
        ka, _, fa, ga = rollout(0)
        out["config"]["memory_chosen_segments"] = {
            "checkpoint_every": ka, "forward_ms": round(fa, 4), "backward_ms": round(ga["ms"], 4),
            "fwd_adjoint_over_plain": round((fa + ga["ms"]) / plain_ms, 3), "tape_and_checkpoint_bytes": ga["bytes"]}
    ps.close()
    return out


def extras(args):
    """Summaries of the other configurations, carried in the default line."""
    ex = {}
    t0 = time.time()
    progress("extras: ref_beach")
    r = plain_bench(args, "ref_beach", 200, 20, extras=False, cpu=False)
    ex["ref_beach"] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "steps": 200, "warmup": 20,
                       "particles": r["config"]["particles"], "vs_rtx4090_published": r["vs_baseline"],
                       "published": "160 fps = 6.4e8 particle-steps/s on an RTX 4090 (the reference's README.md:12-19)",
                       "roofline": {k: r["roofline"][k] for k in ("bound", "frac", "kernel", "kernel_ms")}}
    progress("extras: batch64")
    b = plain_bench(args, "batch64", 100, args.warmup, extras=False, cpu=False)
    ex["batch64"] = {"value": b["value"], "ms_per_step": b["ms_per_step"], "steps": 100, "warmup": args.warmup,
                     "particles": b["config"]["particles"],
                     "single_scene": b["config"].get("single_scene"),
                     "roofline": {k: b["roofline"][k] for k in ("bound", "frac", "kernel", "kernel_ms")}}
    for w in ("adjoint_262k", "adjoint_batch32"):
        progress(f"extras: {w}")
        a = adjoint_bench(args, w, auto=False)
        c = a["config"]
        ex[w] = {"value": a["value"], "steps": a["steps"], "checkpoint_every": c["checkpoint_every"],
                 "particles": c["particles"], "plain_ms": c["plain_ms"], "forward_ms": c["forward_ms"],
                 "backward_ms": c["backward_ms"], "fwd_adjoint_over_plain": c["fwd_adjoint_over_plain"]}
    ex["seconds"] = round(time.time() - t0, 1)
    return ex


# ------------------------------------------------------------ trace summaries

def _rows(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def _base(name):
    return name.split("(")[0].strip()


def _window(rows):
    """Dispatch ids strictly between the fxMark(1) and fxMark(2) dispatches."""
    marks = sorted({int(r["Dispatch_Id"]) for r in rows if _base(r["Kernel_Name"]) == "fxMark"})
    if len(marks) < 2:
        raise SystemExit("summarize: the trace holds fewer than two fxMark dispatches")
    return marks[0], marks[1]


def summarize(args):
    """Reduces rocprofv3 CSV output of one bench.py command to its timed window.

    --trace: kernel_trace.csv of `rocprofv3 --kernel-trace --stats -f csv`;
    --fetch / --write: counter_collection.csv of `--pmc FETCH_SIZE` and
    `--pmc WRITE_SIZE` runs of the same command; --line: that command's JSON
    line.  Prints a markdown summary (whole run and timed window per kernel,
    HBM MB per launch scaled by the copy calibration) and writes --json with
    bytes per launch and the window, which bench.py then attaches to lines of
    the same window."""
    with open(args.line) as fh:
        line = json.loads([ln for ln in fh if ln.strip().startswith("{")][-1])
    win = {k: line["config"]["timed_window"][k] for k in ("steps", "warmup", "autotuned", "graph")}
    K = win["steps"]
    out = []
    dur = {}
    if args.trace:
        rows = _rows(args.trace)
        a, b = _window(rows)
        whole = {}
        for r in rows:
            k = _base(r["Kernel_Name"])
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
            whole.setdefault(k, []).append(d)
            if a < int(r["Dispatch_Id"]) < b:
                dur.setdefault(k, []).append(d)
        tot = sum(sum(v) for v in whole.values())
        out.append("Whole run (every dispatch of the command):\n")
        out.append("| kernel | calls | avg us | min us | max us | share |\n|---|---:|---:|---:|---:|---:|")
        for k, v in sorted(whole.items(), key=lambda kv: -sum(kv[1])):
            if sum(v) / tot < 0.0005:
                continue
            out.append(f"| {k} | {len(v)} | {np.mean(v):.1f} | {min(v):.1f} | {max(v):.1f} | {100 * sum(v) / tot:.1f}% |")
        out.append(f"\nTimed window (dispatches between fxMark(1) and fxMark(2), ids {a}..{b}; {K} steps):\n")
        out.append("| kernel | dispatches | avg us | min us | max us |\n|---|---:|---:|---:|---:|")
        for k, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
            out.append(f"| {k} | {len(v)} | {np.mean(v):.1f} | {min(v):.1f} | {max(v):.1f} |")
        step_us = sum(sum(v) for v in dur.values()) / K
        out.append(f"\nKernel time per step in the window: {step_us:.1f} us; the line's ms_per_step: "
                   f"{1000 * line['ms_per_step']:.1f} us (events, launch gaps included).")
    traffic = {}
    if args.fetch and args.write:
        per = {}
        for path, ctr, fac in ((args.fetch, "FETCH_SIZE", args.fetch_factor), (args.write, "WRITE_SIZE", args.write_factor)):
            rr = _rows(path)
            a2, b2 = _window(rr)
            for r in rr:
                if r.get("Counter_Name", ctr) != ctr or not (a2 < int(r["Dispatch_Id"]) < b2):
                    continue
                per.setdefault(_base(r["Kernel_Name"]), {}).setdefault(ctr, []).append(float(r["Counter_Value"]) * fac)
        out.append(f"\nHBM traffic per launch in the window (FETCH_SIZE x{args.fetch_factor:g} + WRITE_SIZE "
                   f"x{args.write_factor:g}, KiB per dispatch from the copy calibration):\n")
        out.append("| kernel | dispatches | FETCH KiB (scaled) | WRITE KiB (scaled) | HBM MB per launch |\n"
                   "|---|---:|---:|---:|---:|")
        for k, v in sorted(per.items()):
            if "FETCH_SIZE" not in v or "WRITE_SIZE" not in v:
                continue
            f, w = np.mean(v["FETCH_SIZE"]), np.mean(v["WRITE_SIZE"])
            traffic[k] = (f + w) * 1024.0
            out.append(f"| {k} | {len(v['FETCH_SIZE'])} | {f:.0f} | {w:.0f} | {traffic[k] / 1e6:.1f} |")
    limits, vdur = {}, {}
    if args.counters:
        # any other counters: mean per dispatch over the window, per kernel
        per = {}
        for path in args.counters:
            rr = _rows(path)
            a2, b2 = _window(rr)
            for r in rr:
                if a2 < int(r["Dispatch_Id"]) < b2:
                    per.setdefault(_base(r["Kernel_Name"]), {}).setdefault(r["Counter_Name"], []).append(
                        float(r["Counter_Value"]))
        # busy fraction of the texture-data path and VALU issue rate per SIMD
        for path in args.counters:
            rr = _rows(path)
            a2, b2 = _window(rr)
            for r in rr:
                if r.get("Counter_Name") == "SQ_INSTS_VALU" and a2 < int(r["Dispatch_Id"]) < b2:
                    vdur.setdefault(_base(r["Kernel_Name"]), []).append(
                        (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
        for k, v in per.items():
            lim = {}
            if "TD_TD_BUSY_sum" in v and "GRBM_GUI_ACTIVE" in v:
                lim["td_busy"] = round(float(np.mean(v["TD_TD_BUSY_sum"]) / NUM_CU / (np.mean(v["GRBM_GUI_ACTIVE"]) / NUM_XCD)), 4)
            if "SQ_INSTS_VALU" in v and vdur.get(k):
                lim["valu_per_simd_s"] = float(f"{np.mean(v['SQ_INSTS_VALU']) / NUM_SIMD / np.mean(vdur[k]):.4g}")
            if lim:
                limits[k] = lim
        names = sorted({c for v in per.values() for c in v})
        out.append("\nCounters per dispatch in the window (mean):\n")
        out.append("| kernel | dispatches | " + " | ".join(names) + " |\n|---|---:|" + "---:|" * len(names))
        for k, v in sorted(per.items()):
            nd = max(len(x) for x in v.values())
            out.append(f"| {k} | {nd} | " + " | ".join(f"{np.mean(v[c]):.4g}" if c in v else "" for c in names) + " |")
    print("\n".join(out))
    if args.json:
        with open(args.json, "w") as fh:
            cfg = line["config"]
            names = BATCH_KERNEL if cfg["workload"].startswith("batch") else STAGE_KERNEL
            launch = {names[st]: [cfg["stage_threads"][i], cfg.get("stage_ppt", [1] * 6)[i]]
                      for i, st in enumerate(STAGES)}
            json.dump({"workload": cfg["workload"].split(":")[0], "window": win, "launch": launch,
                       "bytes_per_launch": traffic, "limits": limits,
                       "window_avg_us": {k: float(np.mean(v)) for k, v in dur.items()},
                       "fetch_factor": args.fetch_factor, "write_factor": args.write_factor,
                       "line_value": line["value"], "line_ms_per_step": line["ms_per_step"]}, fh, indent=1)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (200; 100 for batch64)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dambreak_4m",
                    choices=["dambreak_16k", "dambreak_4m", "dambreak_4m_rest", "batch64", "adjoint_262k",
                             "adjoint_batch32", "ref_beach"])
    ap.add_argument("--adjoint-steps", type=int, default=20,
                    help="dambreak_4m: steps of the forward+adjoint rollout timed beside the plain step (0 = skip)")
    ap.add_argument("--checkpoint-every", type=int, default=None,
                    help="adjoint segment length (default: 10 for adjoint_262k, 8 for adjoint_batch32, as those "
                         "workloads specify; 0 for the 4M adjoint ratio); 0 = the fewest equal segments whose "
                         "tape fits in 60%% of free device memory")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default block sizes instead of tuning them after the warm-up")
    ap.add_argument("--code-object", default=None, help="step code object to load (default: the built one)")
    ap.add_argument("--stage-threads", default="",
                    help="threads per block of the six stages, comma separated (insert,prefix,sort,pressure,force,"
                         "advance); empty = the library's defaults")
    ap.add_argument("--no-graph", action="store_true", help="launch the timed steps eagerly")
    ap.add_argument("--persistent", default="off", choices=["on", "off"],
                    help="batch stages as persistent launches (per-XCD work queues) or one block per item")
    ap.add_argument("--stats-every", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="oracle steps timed for the all-core CPU baseline; 0 = as many as fill --cpu-seconds (at most 50)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every core this process may use")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-single", action="store_true", help="skip the single-thread CPU figures")
    ap.add_argument("--no-single", action="store_true", help="batch64: skip the one-scene-at-a-time comparison")
    ap.add_argument("--no-extras", action="store_true",
                    help="dambreak_4m: leave out the configured window and the batch64 / adjoint summaries")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (reference for the FETCH_SIZE / WRITE_SIZE "
                         "counters of a kernel of known traffic) and exit")
    ap.add_argument("--summarize", action="store_true", help="reduce rocprofv3 CSVs of a bench.py run (see summarize)")
    ap.add_argument("--line", help="--summarize: the bench.py JSON line of the profiled command")
    ap.add_argument("--trace", help="--summarize: kernel_trace.csv (optional)")
    ap.add_argument("--fetch", help="--summarize: counter_collection.csv of a FETCH_SIZE run")
    ap.add_argument("--write", help="--summarize: counter_collection.csv of a WRITE_SIZE run")
    ap.add_argument("--fetch-factor", type=float, default=2.0)
    ap.add_argument("--write-factor", type=float, default=1.0)
    ap.add_argument("--json", help="--summarize: traffic summary to write")
    ap.add_argument("--counters", nargs="*", default=[],
                    help="--summarize: counter_collection.csv files of other --pmc runs of the same command")
    args = ap.parse_args()

    if args.summarize:
        return summarize(args)

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "copies": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload in ("adjoint_262k", "adjoint_batch32"):
        print(json.dumps(adjoint_bench(args, args.workload, T=args.steps, K=args.checkpoint_every)))
        return

    batch = args.workload == "batch64"
    K = args.steps if args.steps is not None else (100 if batch else 200)
    out = plain_bench(args, args.workload, K, args.warmup, extras=not args.no_extras)
    if args.workload == "dambreak_4m" and not args.no_extras:
        out["extras"] = extras(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
