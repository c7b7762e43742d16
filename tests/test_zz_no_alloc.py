"""fx_scan_device_noalloc, the entry point that allocates nothing, against
fx_scan (the same scan kernels through temporary device buffers).

Called through the C ABI with device buffers the test owns, allocated with the
HIP runtime libfluidsx itself links (libamdhip64.so.7 of ROCm; torch carries a
runtime of its own, which cannot share the device with it in one process).
Two calls in a row on one system, the second with a smaller input, each
bit-identical to fx_scan and to numpy's exclusive cumulative sum; sizes beyond
the scratch array are refused before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

from fluids3_amd.particles import FluidsError, Particles, lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 1 << 27  # FX_SCAN_NOALLOC_MAX in include/fluidsx.h


def hip():
    h = C.CDLL("libamdhip64.so.7")  # already loaded by libfluidsx: the same runtime
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def exclusive(a):
    out = np.zeros_like(a)
    np.cumsum(a[:-1], out=out[1:], dtype=np.uint32)
    return out


def noalloc_scan(ps, values):
    h = hip()
    n = values.shape[0]
    d_in, d_out = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_in), 4 * n) == 0 and h.hipMalloc(C.byref(d_out), 4 * n) == 0
    try:
        v = np.ascontiguousarray(values, np.uint32)
        assert h.hipMemcpy(d_in, v.ctypes.data, 4 * n, H2D) == 0
        assert h.hipDeviceSynchronize() == 0
        st = lib().fx_scan_device_noalloc(ps._h, d_in, d_out, n)
        if st != 0:
            raise FluidsError(f"fx_scan_device_noalloc failed ({st}): {lib().fx_last_error().decode()}")
        out = np.empty(n, np.uint32)
        assert h.hipMemcpy(out.ctypes.data, d_out, 4 * n, D2H) == 0
        return out
    finally:
        h.hipFree(d_in)
        h.hipFree(d_out)


def test_noalloc_scan_matches_fx_scan_twice():
    rng = np.random.default_rng(7)
    with Particles() as ps:
        # three scan levels (more than 1024 level-2 blocks), then one level
        for n in (3 * 1024 * 1024 + 7, 5000):
            print(f"noalloc scan: n = {n}", flush=True)
            v = rng.integers(0, 16, size=n, dtype=np.uint32)
            want = ps.Scan(v)
            got = noalloc_scan(ps, v)
            assert np.array_equal(got, want), f"n={n}: differs from fx_scan"
            assert np.array_equal(got, exclusive(v)), f"n={n}: differs from the exclusive cumulative sum"


def test_noalloc_scan_refuses_sizes_beyond_its_scratch():
    with Particles() as ps:
        st = lib().fx_scan_device_noalloc(ps._h, C.c_void_p(16), C.c_void_p(32), NOALLOC_MAX + 1)
        assert st == -4  # FX_ERANGE, before any launch
        assert b"exceed" in lib().fx_last_error()
