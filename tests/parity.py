"""Comparison helpers shared by the parity tests.

Tolerance definition (density, force and the integrated state):
  rel = |x - x_ref| / max(|x_ref|, tau * rms(x_ref)),  tau = 1e-3,
taken element-wise and reported as its maximum; the floor keeps force
components that cancel to near zero from dominating.  The bar is 1e-5.
Integer outputs (cells, counts, offsets, in-cell ranks, ids) must be equal.
"""
import numpy as np

TAU = 1e-3
TOL = 1e-5


def rel_err(a, b, tau=TAU):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if b.size == 0:
        return 0.0
    floor = tau * np.sqrt(np.mean(b * b))
    den = np.maximum(np.abs(b), floor)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(a - b) / den))


def assert_close(name, a, b, tol=TOL, tau=TAU):
    e = rel_err(a, b, tau)
    assert e <= tol, f"{name}: floored relative error {e:.3e} > {tol:.1e}"
    return e


def assert_int_equal(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{name}: shapes {a.shape} vs {b.shape}"
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{name}: {bad.size} mismatches, first at {bad[:5]}: {a[bad[:5]]} vs {b[bad[:5]]}"


def cell_members_equal(grid_a, grid_b, off, cnt):
    """The member list of every cell holds the same particle indices (order free).
    Entries are compared without the x slab the ranked sort packs above the
    index (bits 28 and up)."""
    mask = np.uint32((1 << 28) - 1)
    grid_a = np.asarray(grid_a, np.uint32) & mask
    grid_b = np.asarray(grid_b, np.uint32) & mask
    n = grid_a.shape[0]
    seg = np.repeat(np.arange(cnt.shape[0]), cnt.astype(np.int64))
    assert seg.shape[0] == n, "counts do not cover the list"
    ka = np.lexsort((grid_a, seg))
    kb = np.lexsort((grid_b, seg))
    return np.array_equal(grid_a[ka], grid_b[kb])


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def valid_mask(state):
    return state["gcell"] != 2147483647
