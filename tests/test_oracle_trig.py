"""The moving wall's sine and cosine (fluids3_amd/csrc/fx_trig.hpp), which the
kernels and the oracle share so that both compute the same bits: against
double-precision sin and cos of the same float argument, within 2 ulps of the
float result, over small, moderate and large arguments, the neighbourhoods of
multiples of pi/2 (where the reduction cancels most digits) and the special
values."""
import math

import numpy as np

from oracle import oracle

F32 = np.float32


def ulps(got, want):
    """|got - want| in units of the float spacing at want (want in float64)."""
    want = np.asarray(want, np.float64)
    spacing = np.spacing(np.abs(want).astype(F32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.spacing(F32(0))))
    return np.abs(np.asarray(got, np.float64) - want) / spacing


def check(x, bound=2.0):
    x = np.asarray(x, F32)
    s, c = oracle.trig(x)
    xs = x.astype(np.float64)
    es, ec = ulps(s, np.sin(xs)), ulps(c, np.cos(xs))
    return float(es.max()), float(ec.max())


def test_moderate_arguments():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-10, 10, 200000), rng.uniform(-200, 200, 200000)])
    es, ec = check(x)
    print(f"|x| <= 200: sin within {es:.2f} ulp, cos within {ec:.2f} ulp")
    assert es <= 2.0 and ec <= 2.0


def test_small_and_large_arguments():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-1e-3, 1e-3, 50000), 10.0 ** rng.uniform(-30, -4, 50000),
                        rng.uniform(-1e4, 1e4, 50000), rng.uniform(-1e6, 1e6, 50000)])
    es, ec = check(x)
    print(f"small and up to 1e6: sin within {es:.2f} ulp, cos within {ec:.2f} ulp")
    assert es <= 2.0 and ec <= 2.0


def test_near_multiples_of_half_pi():
    k = np.arange(-2000, 2001)
    base = (k * (math.pi / 2)).astype(F32)
    x = np.concatenate([np.nextafter(base, F32(-np.inf)), base, np.nextafter(base, F32(np.inf))])
    es, ec = check(x)
    print(f"near k pi/2, |k| <= 2000: sin within {es:.2f} ulp, cos within {ec:.2f} ulp")
    assert es <= 2.0 and ec <= 2.0


def test_special_values():
    s, c = oracle.trig(np.array([0.0, -0.0, np.inf, -np.inf, np.nan], F32))
    assert s[0] == 0.0 and s[1] == 0.0  # of either sign: the reduction adds +0 to -0
    assert c[0] == 1.0 and c[1] == 1.0
    assert np.all(np.isnan(s[2:])) and np.all(np.isnan(c[2:]))
    # exact at the points the series and the reduction make exact
    s, c = oracle.trig(np.array([F32(math.pi / 2)], F32))
    assert s[0] == 1.0
