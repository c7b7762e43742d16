"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product path never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

STAGE_INSERT, STAGE_SCAN, STAGE_SORT, STAGE_PRESSURE, STAGE_FORCE, STAGE_ADVANCE = 1, 2, 4, 8, 16, 32
STAGE_ALL = 63
# oracle stage bits in the order of the device's per-function stages
STAGE_BITS = (STAGE_INSERT, STAGE_SCAN, STAGE_SORT, STAGE_PRESSURE, STAGE_FORCE, STAGE_ADVANCE)

FIELDS = ("pos", "vel", "veval", "force", "press", "clr", "id", "gcell", "gndx", "cnt", "off", "grid")


def build():
    r = subprocess.run(["make", "-B", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building the oracle failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        for name in ("orc_run_f32", "orc_run_f64"):
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = [P, C.c_int] + [P] * 13 + [C.c_int, C.c_int, C.c_int]
        L.orc_time_f32.restype = C.c_double
        L.orc_time_f32.argtypes = [P, C.c_int] + [P] * 13 + [C.c_int, C.c_int]
        L.orc_derive.restype = None
        L.orc_derive.argtypes = [P]
        L.orc_pairs_grid.restype = C.c_int
        L.orc_pairs_grid.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.orc_pairs_brute.restype = C.c_int
        L.orc_pairs_brute.argtypes = [P, C.c_int, P, P, P]
        for name, real in (("f32", C.c_float), ("f64", C.c_double)):
            fn = getattr(L, "orc_grad_" + name)
            fn.restype = C.c_int
            fn.argtypes = [P, C.c_int, P, P, P, P, P, real, C.c_int, C.c_int, C.c_double, C.c_double, P,
                           C.c_int, P, P, P, P, P, P]
            fn = getattr(L, "orc_loss_" + name)
            fn.restype = C.c_double
            fn.argtypes = [P, C.c_int, P, P, P, P, P, real, C.c_int, C.c_int, C.c_double, C.c_double, P,
                           C.c_int]
        L.orc_sample.restype = C.c_int
        L.orc_sample.argtypes = [P, C.c_int, P, P, P, P, P, P, C.c_float, P]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_emit.restype = C.c_int
        L.orc_emit.argtypes = [C.c_int, C.c_float, C.c_int, P, P, P, P]
        L.orc_trig.restype = C.c_int
        L.orc_trig.argtypes = [C.c_int, P, P, P]
        L.orc_grad_mixed.restype = C.c_int
        L.orc_grad_mixed.argtypes = [P, C.c_int, P, P, P, P, P, C.c_float, C.c_int, P, C.c_int, P, P, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data


def make_state(params, pos, vel=None, veval=None, clr=None, ids=None, dtype=np.float32, time=0.0):
    """Initial state dict for n particles (velocities zero unless given)."""
    n = pos.shape[0]
    G = params.gridTotal
    z3 = np.zeros((n, 3), dtype)
    return {
        "pos": np.array(pos, dtype), "vel": z3.copy() if vel is None else np.array(vel, dtype),
        "veval": z3.copy() if veval is None else np.array(veval, dtype), "force": z3.copy(),
        "press": np.zeros(n, dtype), "clr": np.zeros(n, np.uint32) if clr is None else np.array(clr, np.uint32),
        "id": np.arange(n, dtype=np.uint32) if ids is None else np.array(ids, np.uint32),
        "gcell": np.zeros(n, np.int32), "gndx": np.zeros(n, np.int32),
        "cnt": np.zeros(G + 1, np.uint32), "off": np.zeros(G + 1, np.uint32), "grid": np.zeros(n, np.uint32),
        "time": float(time),
    }


def convert(state, dtype):
    out = {}
    for k, v in state.items():
        if k in ("pos", "vel", "veval", "force", "press"):
            out[k] = np.array(v, dtype)
        elif k == "time":
            out[k] = float(v)
        else:
            out[k] = np.array(v, copy=True)
    return out


def run(params, state, nsteps=1, stages=STAGE_ALL, threads=1, fp64=False):
    """Runs the oracle in place on a state dict (see make_state); returns the state."""
    dt = np.float64 if fp64 else np.float32
    s = state
    for k in ("pos", "vel", "veval", "force"):
        s[k] = np.ascontiguousarray(s[k], dtype=dt).reshape(-1, 3)
    s["press"] = np.ascontiguousarray(s["press"], dtype=dt)
    for k in ("clr", "id", "cnt", "off", "grid"):
        s[k] = np.ascontiguousarray(s[k], dtype=np.uint32)
    for k in ("gcell", "gndx"):
        s[k] = np.ascontiguousarray(s[k], dtype=np.int32)
    n = s["pos"].shape[0]
    assert s["cnt"].shape[0] == params.gridTotal + 1
    t = (C.c_double if fp64 else C.c_float)(s["time"])
    fn = lib().orc_run_f64 if fp64 else lib().orc_run_f32
    fn(C.byref(params), n, *[_ptr(s[k]) for k in FIELDS], C.byref(t), nsteps, stages, threads)
    s["time"] = t.value
    return s


def time_steps(params, state, nsteps, threads=1):
    """Seconds for nsteps FP32 oracle steps on `threads` host threads (state advanced in place)."""
    s = run(params, state, 0)  # normalises dtypes
    n = s["pos"].shape[0]
    t = C.c_float(s["time"])
    sec = lib().orc_time_f32(C.byref(params), n, *[_ptr(s[k]) for k in FIELDS], C.byref(t), nsteps, threads)
    s["time"] = t.value
    return sec


def derive(params):
    """Oracle restatement of RebuildAccelGrid + UpdateParams applied to a copy of params."""
    from fluids3_amd.abi import params_copy

    p = params_copy(params)
    lib().orc_derive(C.byref(p))
    return p


def pairs_grid(params, state):
    n = state["pos"].shape[0]
    cand = np.zeros(n, np.int32)
    pairs = np.zeros(n, np.int32)
    pos = np.ascontiguousarray(state["pos"], np.float32)
    gcell = np.ascontiguousarray(state["gcell"], np.int32)
    cnt = np.ascontiguousarray(state["cnt"], np.uint32)
    off = np.ascontiguousarray(state["off"], np.uint32)
    lib().orc_pairs_grid(C.byref(params), n, _ptr(pos), _ptr(gcell), _ptr(cnt), _ptr(off), _ptr(cand), _ptr(pairs))
    return cand, pairs


def pairs_brute(params, pos):
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    pairs = np.zeros(n, np.int32)
    dens = np.zeros(n, np.float64)
    lib().orc_pairs_brute(C.byref(params), n, _ptr(pos), _ptr(pairs), _ptr(dens))
    return pairs, dens


def _initial(state, dt):
    pos = np.ascontiguousarray(state["pos"], dt).reshape(-1, 3)
    n = pos.shape[0]
    vel = np.ascontiguousarray(state.get("vel", np.zeros((n, 3))), dt).reshape(-1, 3)
    veval = np.ascontiguousarray(state.get("veval", vel), dt).reshape(-1, 3)
    clr = np.ascontiguousarray(state.get("clr", np.zeros(n)), np.uint32)
    ids = np.ascontiguousarray(state.get("id", np.arange(n)), np.uint32)
    return n, pos, vel, veval, clr, ids


def _coef(coef):
    return (0, 0.0, 0.0) if coef is None else (1, float(coef[0]), float(coef[1]))


def grad(params, state, nsteps, target, fp64=False, coef=None, threads=1):
    """Loss (1/n) sum |x_T - x*|^2 of an nsteps rollout from `state` (pos, vel,
    veval, clr, id, time) against target (n x 3, by particle id), and its
    gradient with respect to the initial pos, vel and veval (by id), viscosity and
    internal stiffness.  coef = (visc, k) evaluates the force with those
    values in the rollout's precision instead of the FParams_t floats."""
    dt = np.float64 if fp64 else np.float32
    n, pos, vel, veval, clr, ids = _initial(state, dt)
    tgt = np.ascontiguousarray(target, dt).reshape(-1, 3)
    assert tgt.shape[0] == n
    gp = np.zeros((n, 3), dt)
    gv = np.zeros((n, 3), dt)
    ge = np.zeros((n, 3), dt)
    loss, gvisc, gk = C.c_double(), C.c_double(), C.c_double()
    use, visc, k = _coef(coef)
    fn = lib().orc_grad_f64 if fp64 else lib().orc_grad_f32
    fn(C.byref(params), n, _ptr(pos), _ptr(vel), _ptr(veval), _ptr(clr), _ptr(ids), float(state.get("time", 0.0)),
       nsteps, use, visc, k, _ptr(tgt), threads, C.byref(loss), _ptr(gp), _ptr(gv), _ptr(ge), C.byref(gvisc),
       C.byref(gk))
    return {"loss": loss.value, "gpos": gp, "gvel": gv, "gveval": ge, "gvisc": gvisc.value, "gk": gk.value}


def loss(params, state, nsteps, target, fp64=True, coef=None, threads=1):
    """Loss of an nsteps rollout (see grad), no gradient."""
    dt = np.float64 if fp64 else np.float32
    n, pos, vel, veval, clr, ids = _initial(state, dt)
    tgt = np.ascontiguousarray(target, dt).reshape(-1, 3)
    use, visc, k = _coef(coef)
    fn = lib().orc_loss_f64 if fp64 else lib().orc_loss_f32
    return fn(C.byref(params), n, _ptr(pos), _ptr(vel), _ptr(veval), _ptr(clr), _ptr(ids),
              float(state.get("time", 0.0)), nsteps, use, visc, k, _ptr(tgt), threads)


def by_id(values, ids):
    """Reorders per-slot values into particle-id order."""
    out = np.empty_like(values)
    out[np.asarray(ids, np.int64)] = values
    return out


def grad_mixed(params, state, nsteps, target, threads=1):
    """FP64 adjoint over the FP32 rollout's tape (same outputs as grad, float64)."""
    n, pos, vel, veval, clr, ids = _initial(state, np.float32)
    tgt = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    g = {k: np.zeros((n, 3)) for k in ("gpos", "gvel", "gveval")}
    loss, gvisc, gk = C.c_double(), C.c_double(), C.c_double()
    lib().orc_grad_mixed(C.byref(params), n, _ptr(pos), _ptr(vel), _ptr(veval), _ptr(clr), _ptr(ids),
                         float(state.get("time", 0.0)), nsteps, _ptr(tgt), threads, C.byref(loss), _ptr(g["gpos"]),
                         _ptr(g["gvel"]), _ptr(g["gveval"]), C.byref(gvisc), C.byref(gk))
    g.update({"loss": loss.value, "gvisc": gvisc.value, "gk": gk.value})
    return g


def sample(params, state, res, bmin, bmax, scalar=1.0):
    """Oracle sampleParticles on a cell-sorted state: (res_z, res_y, res_x) float32."""
    # every array passed by address is bound to a name for the duration of the call
    pos = np.ascontiguousarray(state["pos"], np.float32)
    cnt = np.ascontiguousarray(state["cnt"], np.uint32)
    off = np.ascontiguousarray(state["off"], np.uint32)
    r = np.ascontiguousarray(res, np.int32)
    lo = np.ascontiguousarray(bmin, np.float32)
    hi = np.ascontiguousarray(bmax, np.float32)
    out = np.zeros((int(r[2]), int(r[1]), int(r[0])), np.float32)
    lib().orc_sample(C.byref(params), pos.shape[0], _ptr(pos), _ptr(cnt), _ptr(off), _ptr(r), _ptr(lo), _ptr(hi),
                     float(scalar), _ptr(out))
    return out


def splitmix64(x):
    """One SplitMix64 step from state x (the hash behind the emitter's draws)."""
    return int(lib().orc_splitmix64(int(x) & 0xFFFFFFFFFFFFFFFF))


def emit(state, frame, count):
    """Oracle emitParticles on copies of pos/vel/veval/force: returns the four (n, 3) float32 arrays."""
    pos = np.array(state["pos"], np.float32, order="C", copy=True)
    n = pos.shape[0]
    vel = np.array(state.get("vel", np.zeros((n, 3))), np.float32, order="C", copy=True)
    veval = np.array(state.get("veval", np.zeros((n, 3))), np.float32, order="C", copy=True)
    force = np.array(state.get("force", np.zeros((n, 3))), np.float32, order="C", copy=True)
    lib().orc_emit(n, float(frame), int(count), _ptr(pos), _ptr(vel), _ptr(veval), _ptr(force))
    return pos, vel, veval, force


def trig(x):
    """The moving wall's sine and cosine (the kernels' own, fluids3_amd/csrc/fx_trig.hpp) of float32 x."""
    x = np.ascontiguousarray(x, np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_trig(x.size, _ptr(x), _ptr(s), _ptr(c))
    return s, c

