"""Pins of the CPU adjoint oracle (oracle/sph_adjoint.hpp).

The FP64 adjoint is checked against FP64 central differences of the FP64
rollout (oracle.loss) in every input kind it differentiates: initial pos, vel
and veval components, viscosity, internal stiffness.  The rollout is piecewise
smooth (neighbour cut-offs, wall bands and limiters switch), so a difference
quotient whose step straddles a switch is detected by comparing two step
sizes and that component is replaced by another seeded one.  A ballistic
particle gives a closed form.  The bar (1e-4 relative against step sizes
of 1e-6) is far below what a dropped or sign-flipped pair term moves.
"""
import numpy as np
import pytest

from fluids3_amd import derive_params, fill_points, scenes
from fluids3_amd.abi import params_copy
from oracle import oracle

N_SMALL = 2000
STEPS = 20
THREADS = 8


@pytest.fixture(scope="module")
def problem():
    sc = scenes.scaled_dam_break(N_SMALL)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    # target: the same rollout at viscosity 0.30 and stiffness 1.2
    P2 = params_copy(P)
    P2.pvisc = 0.30
    P2.pintstiff = 1.2
    P2.vterm = np.float32(P2.lapkern) * np.float32(0.30)
    r = oracle.run(P2, oracle.make_state(P, pos, clr=clr, dtype=np.float64), STEPS, fp64=True, threads=THREADS)
    target = oracle.by_id(r["pos"], r["id"])
    st = {"pos": pos.astype(np.float64), "clr": clr}
    g = oracle.grad(P, st, STEPS, target, fp64=True, coef=(0.35, 1.0), threads=THREADS)
    return P, st, target, g


def _loss(P, st, target, coef=(0.35, 1.0), **fields):
    s = dict(st)
    s.update(fields)
    return oracle.loss(P, s, STEPS, target, fp64=True, coef=coef, threads=THREADS)


def _cd(f, x0, eps):
    return (f(x0 + eps) - f(x0 - eps)) / (2 * eps)


def _smooth_cd(f, x0, eps):
    """Central difference at eps, or None when eps and eps/3 disagree (a switch inside the step)."""
    a, b = _cd(f, x0, eps), _cd(f, x0, eps / 3)
    if abs(a - b) > 1e-6 * max(abs(a), abs(b), 1e-12):
        return None
    return b


def test_adjoint_scalar_parameters_match_central_differences(problem):
    P, st, target, g = problem
    for name, x0, ad in (("visc", 0.35, g["gvisc"]), ("k", 1.0, g["gk"])):
        f = (lambda v: _loss(P, st, target, coef=(v, 1.0))) if name == "visc" else \
            (lambda v: _loss(P, st, target, coef=(0.35, v)))
        cd = None
        for eps in (1e-6, 3e-7, 1e-7):
            cd = _smooth_cd(f, x0, eps * x0)
            if cd is not None:
                break
        assert cd is not None, name
        print(f"d/d{name}: adjoint {ad:.10e}, central difference {cd:.10e}, rel {abs(ad - cd) / abs(cd):.2e}")
        assert abs(ad - cd) <= 1e-4 * abs(cd), (name, ad, cd)


def test_adjoint_velocity_components_match_central_differences(problem):
    P, st, target, g = problem
    n = st["pos"].shape[0]
    rng = np.random.default_rng(5)
    checked = 0
    errs = []
    for c in rng.permutation(n * 3):
        if checked == 8:
            break
        i, a = divmod(int(c), 3)
        for which in ("vel", "veval", "pos"):
            def f(x, which=which):
                arr = np.zeros((n, 3))
                arr[i, a] = x
                if which == "pos":
                    return _loss(P, st, target, pos=st["pos"] + arr)
                other = np.zeros((n, 3))
                return _loss(P, st, target, **{which: arr, ("veval" if which == "vel" else "vel"): other})
            cd = _smooth_cd(f, 0.0, 1e-6)
            if cd is None:
                break
            ad = g["g" + which][i, a]
            errs.append(abs(ad - cd) / max(abs(cd), 1e-3 * np.abs(g["g" + which]).max()))
        else:
            checked += 1
    assert checked == 8
    print(f"pos/vel/veval components: max rel error {max(errs):.2e} over {len(errs)}")
    assert max(errs) <= 1e-4, errs


def test_adjoint_ballistic_particle_closed_form():
    # one particle in the middle of a large box: no neighbours, no wall, no limit;
    # x_T = x0 + (v0 T + g dt (1 + ... + T)) dt/ss, so dL/dv0 = (2/n)(x_T - x*) T dt/ss
    sc = scenes.dam_break(1, (-200, -200, -200), (200, 200, 200), (-1, -1, -1), (1, 1, 1))
    P = derive_params(sc)
    T = 5
    x0 = np.array([[0.5, 40.0, -0.25]])
    v0 = np.array([[0.3, 0.0, -0.2]])
    tgt = np.array([[1.0, 2.0, 3.0]])
    st = {"pos": x0, "vel": v0, "veval": v0}
    for fp64 in (True, False):
        g = oracle.grad(P, st, T, tgt, fp64=fp64)
        dts = P.dt / P.sim_scale
        xT = x0 + (v0 * T + np.array(P.gravity.tolist()) * P.dt * T * (T + 1) / 2) * dts
        expect = 2 * (xT - tgt) * T * dts
        tol = 1e-12 if fp64 else 1e-5
        assert np.allclose(g["gvel"], expect, rtol=tol, atol=0), (fp64, g["gvel"], expect)
        assert np.all(g["gveval"] == 0)
        assert g["gvisc"] == 0 and g["gk"] == 0
        assert g["loss"] == pytest.approx(float(((xT - tgt) ** 2).sum()), rel=1e-12 if fp64 else 1e-5)


def test_adjoint_fp32_close_to_fp64(problem):
    P, st, target, g64 = problem
    st32 = {"pos": st["pos"].astype(np.float32), "clr": st["clr"]}
    g32 = oracle.grad(P, st32, STEPS, target.astype(np.float32), fp64=False, threads=THREADS)
    for k in ("gvel", "gveval"):
        e = np.linalg.norm(g32[k] - g64[k]) / np.linalg.norm(g64[k])
        assert e < 1e-3, (k, e)
    for k in ("gvisc", "gk"):
        assert abs(g32[k] - g64[k]) <= 1e-3 * abs(g64[k]), (k, g32[k], g64[k])


def test_adjoint_zero_steps_is_the_seed():
    sc = scenes.scaled_dam_break(300)
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    tgt = pos.astype(np.float64) + 0.5
    g = oracle.grad(P, {"pos": pos, "clr": clr}, 0, tgt, fp64=True)
    assert np.all(g["gvel"] == 0) and np.all(g["gveval"] == 0)
    assert g["loss"] == pytest.approx(0.75, rel=1e-12)


def test_adjoint_walls_slope_friction_moving_wall():
    """Every wall branch of the integrator's reverse: a sloped floor with
    friction, the moving X- wall (stiffness scaled by the wall's sine, which
    depends on z and on time), X+/Z walls and the ceiling.  A small block is
    launched at the corners so that all bands are entered within the rollout;
    FP64 adjoint against FP64 central differences."""
    sc = scenes.dam_break(600, (-12, 0, -12), (12, 14, 12), (-11.5, 0.5, -11.5), (11.5, 13.5, 11.5))
    sc.bound_slope = 0.12
    sc.bound_friction = 0.3
    sc.bound_wall_force = 3.0
    sc.bound_wall_freq = 2.0
    P = derive_params(sc)
    pos, clr = fill_points(sc)
    n = pos.shape[0]
    rng = np.random.default_rng(9)
    v0 = rng.uniform(-1.0, 1.0, (n, 3))
    T = 12
    tgt = pos.astype(np.float64) + rng.uniform(-1, 1, (n, 3))
    st = {"pos": pos.astype(np.float64), "clr": clr, "vel": v0, "veval": v0}
    g = oracle.grad(P, st, T, tgt, fp64=True, coef=(0.35, 1.0), threads=THREADS)

    def L(**kw):
        s = dict(st)
        s.update(kw)
        return oracle.loss(P, s, T, tgt, fp64=True, coef=(0.35, 1.0), threads=THREADS)

    checked, errs = 0, []
    for c in rng.permutation(n * 3):
        if checked == 10:
            break
        i, a = divmod(int(c), 3)
        row = []
        for which in ("pos", "vel"):
            def f(x, which=which):
                arr = st[which].copy()
                arr[i, a] += x
                return L(**{which: arr, **({"veval": v0} if which == "vel" else {})})
            cd = _smooth_cd(f, 0.0, 1e-6)
            if cd is None:
                break
            ad = g["g" + which][i, a]
            row.append(abs(ad - cd) / max(abs(cd), 1e-3 * np.abs(g["g" + which]).max()))
        else:
            errs += row
            checked += 1
    assert checked == 10
    print(f"walls: max rel error {max(errs):.2e} over {len(errs)} components")
    assert max(errs) <= 1e-4, errs
