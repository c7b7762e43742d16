"""Python mirror of the reference's Particles interface over the C ABI (libfluidsx.so).

Method names follow fluids5.0/src/particles.h:35-70 (Restart, Run,
InsertParticles, PrefixScanParticles, CountingSort, ComputePressure,
ComputeForce, Advance, Retrieve ...).  Every call goes through the C ABI of
include/fluidsx.h and runs the gfx950 code object; there is no CPU fallback:
a missing library or a failing call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfluidsx.so")
CODE_OBJECT = os.path.join(_HERE, "particles_gfx950.hsaco")

_lib = None

# every symbol include/fluidsx.h declares: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "fx_scene_derive": (C.c_int, [C.POINTER(abi.Scene), C.POINTER(abi.FParams)]),
    "fx_scene_spacing": (C.c_float, [C.POINTER(abi.Scene)]),
    "fx_scene_fill": (C.c_int, [C.POINTER(abi.Scene), _P, _P, C.c_int32]),
    "fx_create": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "fx_destroy": (None, [_P]),
    "fx_restart": (C.c_int, [_P, C.POINTER(abi.Scene)]),
    "fx_restart_batch": (C.c_int, [_P, C.POINTER(abi.Scene), C.c_int32]),
    "fx_num_scenes": (C.c_int32, [_P]),
    "fx_scene_range": (C.c_int, [_P, C.c_int32] + [C.POINTER(C.c_int32)] * 4),
    "fx_scene_params": (C.c_int, [_P, C.c_int32, C.POINTER(abi.FParams)]),
    "fx_get_params": (C.c_int, [_P, C.POINTER(abi.FParams)]),
    "fx_set_params": (C.c_int, [_P, C.POINTER(abi.FParams)]),
    "fx_num_points": (C.c_int32, [_P]),
    "fx_grid_total": (C.c_int32, [_P]),
    "fx_cell_table_size": (C.c_int32, [_P]),
    "fx_buffer_stride": (C.c_int32, [C.c_int32]),
    "fx_upload": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "fx_download": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "fx_download_accel": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "fx_sample_particles": (C.c_int, [_P, _P, _P, _P, _P, C.c_float]),
    "fx_save_points": (C.c_int, [_P, C.c_char_p]),
    "fx_load_points": (C.c_int, [_P, C.c_char_p]),
    "fx_stage": (C.c_int, [_P, C.c_int32]),
    "fx_run": (C.c_int, [_P, C.c_int32]),
    "fx_synchronize": (C.c_int, [_P]),
    "fx_mark": (C.c_int, [_P, C.c_int32]),
    "fx_scan": (C.c_int, [_P, _P, _P, C.c_int32]),
    "fx_scan_device_noalloc": (C.c_int, [_P, _P, _P, C.c_int32]),
    "fx_graph_capture": (C.c_int, [_P]),
    "fx_graph_launch": (C.c_int, [_P, C.c_int32]),
    "fx_profile_stages": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "fx_time_steps": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "fx_bench_steps": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fx_count_pairs": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fx_set_block_size": (C.c_int, [_P, C.c_int32]),
    "fx_set_stage_block_size": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "fx_stage_block_size": (C.c_int32, [_P, C.c_int32]),
    "fx_autotune": (C.c_int, [_P, C.c_int32, _P, _P]),
    "fx_set_stage_ppt": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "fx_stage_ppt": (C.c_int32, [_P, C.c_int32]),
    "fx_set_persistent": (C.c_int, [_P, C.c_int32]),
    "fx_adjoint_configure": (C.c_int, [_P, C.c_char_p, C.c_int32, C.c_int32]),
    "fx_adjoint_set_target": (C.c_int, [_P, _P, C.c_int64]),
    "fx_adjoint_forward": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "fx_adjoint_backward": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
    "fx_adjoint_memory": (C.c_int64, [_P]),
    "fx_adjoint_segment": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "fx_emit_particles": (C.c_int, [_P, C.c_float, C.c_int32]),
    "fx_scene_from_xml": (C.c_int, [C.c_char_p, C.c_int32, _P, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32]),
    "fx_scene_xml_ids": (C.c_int, [C.c_char_p, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "fx_last_error": (C.c_char_p, []),
    "fx_version": (C.c_char_p, []),
}


class FluidsError(RuntimeError):
    pass


def lib():
    """The loaded libfluidsx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FluidsError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, what):
    if status != 0:
        msg = lib().fx_last_error().decode(errors="replace")
        raise FluidsError(f"{what} failed ({status}): {msg}")


def derive_params(scene):
    p = abi.FParams()
    check(lib().fx_scene_derive(C.byref(scene), C.byref(p)), "fx_scene_derive")
    return p


def scene_spacing(scene):
    return float(lib().fx_scene_spacing(C.byref(scene)))


def fill_points(scene, n=None):
    """Initial lattice of a scene on the host: (pos float32[n,3], clr uint32[n])."""
    n = scene.num_points if n is None else n
    pos = np.zeros((n, 3), np.float32)
    clr = np.zeros(n, np.uint32)
    got = lib().fx_scene_fill(C.byref(scene), pos.ctypes.data, clr.ctypes.data, n)
    if got < 0:
        check(got, "fx_scene_fill")
    return pos[:got], clr[:got]


class Particles:
    """One simulation on the GPU: the reference's Particles (particles.h:29-96)."""

    def __init__(self, code_object=CODE_OBJECT):
        if not os.path.exists(code_object):
            raise FluidsError(f"{code_object} is missing: build it with __graft_entry__.build()")
        h = C.c_void_p()
        check(lib().fx_create(code_object.encode(), C.byref(h)), "fx_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- set-up
    def Restart(self, scene):
        check(lib().fx_restart(self._h, C.byref(scene)), "fx_restart")

    def RestartBatch(self, scene_list):
        """Many independent scenes in one system, each stage one launch over all."""
        arr = (abi.Scene * len(scene_list))(*scene_list)
        check(lib().fx_restart_batch(self._h, arr, len(scene_list)), "fx_restart_batch")

    def NumScenes(self):
        return int(lib().fx_num_scenes(self._h))

    def SceneRange(self, s):
        """(first particle, particles, first cell, cells incl. the tail slot) of scene s."""
        v = [C.c_int32() for _ in range(4)]
        check(lib().fx_scene_range(self._h, s, *[C.byref(x) for x in v]), "fx_scene_range")
        return tuple(x.value for x in v)

    def SceneParams(self, s):
        p = abi.FParams()
        check(lib().fx_scene_params(self._h, s, C.byref(p)), "fx_scene_params")
        return p

    @property
    def params(self):
        p = abi.FParams()
        check(lib().fx_get_params(self._h, C.byref(p)), "fx_get_params")
        return p

    def UpdateParams(self, p):
        check(lib().fx_set_params(self._h, C.byref(p)), "fx_set_params")

    def NumPoints(self):
        return int(lib().fx_num_points(self._h))

    def GridTotal(self):
        return int(lib().fx_grid_total(self._h))

    def CellTableSize(self):
        return int(lib().fx_cell_table_size(self._h))

    def SetBlockSize(self, threads):
        check(lib().fx_set_block_size(self._h, threads), "fx_set_block_size")

    def SetStageBlockSize(self, stage, threads):
        check(lib().fx_set_stage_block_size(self._h, stage, threads), "fx_set_stage_block_size")

    def StageBlockSize(self, stage):
        return int(lib().fx_stage_block_size(self._h, stage))

    # ---- stages (particles.cpp:382-394)
    def InsertParticles(self):
        check(lib().fx_stage(self._h, abi.STAGE_INSERT), "insert")

    def PrefixScanParticles(self):
        check(lib().fx_stage(self._h, abi.STAGE_PREFIX), "prefix")

    def CountingSort(self):
        check(lib().fx_stage(self._h, abi.STAGE_SORT), "sort")

    def ComputePressure(self):
        check(lib().fx_stage(self._h, abi.STAGE_PRESSURE), "pressure")

    def ComputeForce(self):
        check(lib().fx_stage(self._h, abi.STAGE_FORCE), "force")

    def Advance(self):
        """Advance and AdvanceTime (particles.cpp:413-418, 396-400)."""
        check(lib().fx_stage(self._h, abi.STAGE_ADVANCE), "advance")

    def Stage(self, k):
        check(lib().fx_stage(self._h, k), f"stage {k}")

    def Run(self, nsteps=1):
        check(lib().fx_run(self._h, nsteps), "fx_run")

    def Synchronize(self):
        check(lib().fx_synchronize(self._h), "fx_synchronize")

    def Mark(self, tag):
        """One dispatch of the no-op kernel fxMark(tag), marking this point of the stream in traces."""
        check(lib().fx_mark(self._h, int(tag)), "fx_mark")

    def Scan(self, values):
        """Exclusive prefix sum of a uint32 array with the grid-offset scan kernels."""
        a = np.ascontiguousarray(values, dtype=np.uint32)
        out = np.empty_like(a)
        check(lib().fx_scan(self._h, a.ctypes.data, out.ctypes.data, a.shape[0]), "fx_scan")
        return out

    def CaptureGraph(self):
        check(lib().fx_graph_capture(self._h), "fx_graph_capture")

    def LaunchGraph(self, nlaunch):
        check(lib().fx_graph_launch(self._h, nlaunch), "fx_graph_launch")

    # ---- transfers
    def Retrieve(self, buf, n=None):
        """Device copy of particle buffer `buf` (FPOS..FID) as a numpy array."""
        n = self.NumPoints() if n is None else n
        dt, comp = abi.PBUF_LAYOUT[buf]
        out = np.empty((n, comp) if comp > 1 else (n,), dtype=dt)
        check(lib().fx_download(self._h, buf, out.ctypes.data, n), f"download {buf}")
        return out

    def Commit(self, buf, arr):
        dt, comp = abi.PBUF_LAYOUT[buf]
        a = np.ascontiguousarray(arr, dtype=dt)
        n = a.shape[0]
        if comp > 1 and a.shape[1:] != (comp,):
            raise ValueError(f"buffer {buf} expects shape (n, {comp})")
        check(lib().fx_upload(self._h, buf, a.ctypes.data, n), f"upload {buf}")

    def SampleParticles(self, res, bmin, bmax, scalar=1.0):
        """Gaussian splat of the particles on a brick (sampleParticles); returns a
        float32 array of shape (res_z, res_y, res_x)."""
        r = np.ascontiguousarray(res, np.int32)
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        out = np.empty((int(r[2]), int(r[1]), int(r[0])), np.float32)
        check(lib().fx_sample_particles(self._h, out.ctypes.data, r.ctypes.data, lo.ctypes.data, hi.ctypes.data,
                                        float(scalar)), "fx_sample_particles")
        return out

    def EmitParticles(self, frame, emit):
        """Respawns min(emit, n) particles in the emitter box (emitParticles); frame seeds the draw."""
        check(lib().fx_emit_particles(self._h, float(frame), int(emit)), "fx_emit_particles")

    def SavePoints(self, path):
        """Writes the particle state as a .pts file (Fluids 4.0 layout plus veval and id)."""
        check(lib().fx_save_points(self._h, os.fsencode(path)), "fx_save_points")

    def LoadPoints(self, path):
        """Reads a .pts file written by SavePoints or by Fluids 4.0 (same particle count)."""
        check(lib().fx_load_points(self._h, os.fsencode(path)), "fx_load_points")

    def RetrieveAccel(self, buf, n):
        out = np.empty(n, dtype=np.uint32)
        check(lib().fx_download_accel(self._h, buf, out.ctypes.data, n), f"download_accel {buf}")
        return out

    def state(self):
        """All particle buffers and the grid counts/offsets, as numpy arrays."""
        ncell = self.CellTableSize()
        st = {name: self.Retrieve(b) for name, b in (
            ("pos", abi.FPOS), ("vel", abi.FVEL), ("veval", abi.FVEVAL), ("force", abi.FFORCE),
            ("press", abi.FPRESS), ("clr", abi.FCLR), ("gcell", abi.FGCELL), ("gndx", abi.FGNDX),
            ("id", abi.FID))}
        st["cnt"] = self.RetrieveAccel(abi.AGRIDCNT, ncell)
        st["off"] = self.RetrieveAccel(abi.AGRIDOFF, ncell)
        st["grid"] = self.RetrieveAccel(abi.AGRID, self.NumPoints())
        st["time"] = self.params.time
        return st

    # ---- timing
    def ProfileStages(self, nsteps):
        out = (C.c_double * (abi.NUM_STAGES + 1))()
        check(lib().fx_profile_stages(self._h, nsteps, out), "fx_profile_stages")
        return {**{abi.STAGE_NAMES[k]: out[k] for k in range(abi.NUM_STAGES)}, "step": out[abi.NUM_STAGES]}

    def BenchSteps(self, nsteps, use_graph=True):
        """(total ms, {stage: mean ms}) of nsteps steps timed with HIP events on the library stream."""
        tot = C.c_double()
        st = (C.c_double * abi.NUM_STAGES)()
        check(lib().fx_bench_steps(self._h, nsteps, 1 if use_graph else 0, C.byref(tot), st), "fx_bench_steps")
        return tot.value, {abi.STAGE_NAMES[k]: st[k] for k in range(abi.NUM_STAGES)}

    def SetStagePpt(self, stage, ppt):
        check(lib().fx_set_stage_ppt(self._h, stage, ppt), "fx_set_stage_ppt")

    def StagePpt(self, stage):
        return lib().fx_stage_ppt(self._h, stage)

    def SetPersistent(self, on=True):
        """Persistent (work-queue) launches of the batch stages."""
        check(lib().fx_set_persistent(self._h, 1 if on else 0), "fx_set_persistent")

    def CountPairs(self):
        """(candidates, pairs) summed over particles on the current cell-sorted state."""
        c, p = C.c_int64(), C.c_int64()
        check(lib().fx_count_pairs(self._h, C.byref(c), C.byref(p)), "fx_count_pairs")
        return c.value, p.value

    def Autotune(self, nsteps=5):
        """Picks each stage's block size (fx_autotune); returns {stage: (threads, ms)}."""
        th = (C.c_int32 * abi.NUM_STAGES)()
        ms = (C.c_double * abi.NUM_STAGES)()
        check(lib().fx_autotune(self._h, nsteps, th, ms), "fx_autotune")
        return {abi.STAGE_NAMES[k]: (th[k], ms[k]) for k in range(abi.NUM_STAGES)}

    # ---- reverse mode
    def AdjointConfigure(self, nsteps, every=0, code_object=None):
        """Checkpoints every `every` steps of an nsteps rollout from the current state
        (every <= 0: the fewest equal segments that fit in 60 % of the free device memory)."""
        path = None if code_object is None else code_object.encode()
        check(lib().fx_adjoint_configure(self._h, path, nsteps, every), "fx_adjoint_configure")

    def AdjointSegment(self):
        """Segment length in use (the chosen one after AdjointConfigure(n, 0))."""
        e = C.c_int32()
        check(lib().fx_adjoint_segment(self._h, C.byref(e)), "fx_adjoint_segment")
        return e.value

    def AdjointSetTarget(self, target):
        """Target positions, (n, 3) float32 in particle-id order."""
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, 3)
        check(lib().fx_adjoint_set_target(self._h, t.ctypes.data, t.shape[0]), "fx_adjoint_set_target")

    def AdjointForward(self):
        """(loss per scene as float64 array, GPU ms) of the configured rollout."""
        loss = (C.c_double * self.NumScenes())()
        ms = C.c_double()
        check(lib().fx_adjoint_forward(self._h, loss, C.byref(ms)), "fx_adjoint_forward")
        return np.array(loss[:], dtype=np.float64), ms.value

    def AdjointBackward(self):
        """Gradients of the last AdjointForward's loss: dict with gpos, gvel, gveval ((n, 3) float32 by
        particle id), gvisc, gk (float64 per scene) and ms (GPU time)."""
        n, ns = self.NumPoints(), self.NumScenes()
        gp = np.empty((n, 3), np.float32)
        gv = np.empty((n, 3), np.float32)
        ge = np.empty((n, 3), np.float32)
        gvisc = (C.c_double * ns)()
        gk = (C.c_double * ns)()
        ms = C.c_double()
        check(lib().fx_adjoint_backward(self._h, gp.ctypes.data, gv.ctypes.data, ge.ctypes.data, gvisc, gk,
                                        C.byref(ms)), "fx_adjoint_backward")
        return {"gpos": gp, "gvel": gv, "gveval": ge, "gvisc": np.array(gvisc[:]), "gk": np.array(gk[:]),
                "ms": ms.value}

    def AdjointMemory(self):
        return lib().fx_adjoint_memory(self._h)

    def TimeSteps(self, nsteps, use_graph=True):
        ms = C.c_double()
        check(lib().fx_time_steps(self._h, nsteps, 1 if use_graph else 0, C.byref(ms)), "fx_time_steps")
        return ms.value
