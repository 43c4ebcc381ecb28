"""st_amd_element_stats_noalloc (include/stapi_amd.h), the element
statistics for callers that cannot allocate device memory, against
st_amd_element_stats on the same resident records, through the C-ABI: two
calls in a row on one context, the second with a smaller grid (the shared
scratch keeps the first call's counts, which must not leak into the
second), both with fixed and autoscaled ranges; grids equal bin for bin and
statistics equal bit for bit (the same kernels and host arithmetic run on
different buffers).  A grid beyond the scratch's 1,048,576 bins is refused
with a message, as is a missing grid."""
import ctypes as C

import numpy as np
import pytest

from flux_stats import element_stats
from soltrace_amd import stapi, workloads

pytestmark = pytest.mark.gpu


def _both(lib, ctx, stage, elem, nx, ny, autoscale, absorbed, bounds):
    a = element_stats(lib, ctx, "st_amd_element_stats", stage, elem, nx, ny, autoscale, absorbed, 950.0, bounds)
    b = element_stats(lib, ctx, "st_amd_element_stats_noalloc", stage, elem, nx, ny, autoscale, absorbed, 950.0,
                      bounds)
    assert a is not None and b is not None
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert a[2] == b[2]
    return a


@pytest.mark.parametrize("case", ["tower_flat", "trough_cylinder"])
def test_noalloc_equals_allocating_entry_point(hip_lib, case):
    if case == "tower_flat":
        sc, half = workloads.tower(n_helio=100, raycount=20000, r_max=500.0), (-7.5, 7.5, -7.5, 7.5)
    else:
        sc, half = workloads.trough(raycount=20000), (-1.0, 1.0, -6.0, 6.0)
    print(f"noalloc {case}", flush=True)
    ctx = hip_lib.st_create_context()
    try:
        sc.apply(hip_lib, ctx)
        assert stapi.run(hip_lib, ctx, 21) > 0, stapi.messages(hip_lib, ctx)
        big = _both(hip_lib, ctx, 1, 0, 200, 150, 0, 1, half)
        assert big[1].sum() > 1000
        small = _both(hip_lib, ctx, 1, 0, 30, 20, 0, 1, half)
        assert small[1].sum() == big[1].sum() or case == "trough_cylinder"
        _both(hip_lib, ctx, 1, 0, 64, 64, 1, 0, half)
        _both(hip_lib, ctx, 1, 0, 10, 12, 1, 0, half)
    finally:
        hip_lib.st_free_context(ctx)


def test_noalloc_refuses_what_it_cannot_hold(hip_lib):
    sc = workloads.trough(raycount=2000)
    ctx = hip_lib.st_create_context()
    try:
        sc.apply(hip_lib, ctx)
        assert stapi.run(hip_lib, ctx, 3) > 0
        f = hip_lib.dll.st_amd_element_stats_noalloc
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                      C.POINTER(C.c_double), C.POINTER(C.c_uint), C.POINTER(C.c_double)]
        f.restype = C.c_int
        b = (C.c_double * 4)(-1.0, 1.0, -6.0, 6.0)
        st = (C.c_double * 15)()
        n_before = hip_lib.st_num_messages(ctx)
        grid = (C.c_uint * 4)()
        assert f(ctx, 1, 0, 1025, 1024, 0, 1, 950.0, b, grid, st) == -1  # 1,049,600 bins
        assert hip_lib.st_num_messages(ctx) == n_before + 1
        assert "exceed" in stapi.messages(hip_lib, ctx)[-1]
        assert f(ctx, 1, 0, 10, 10, 0, 1, 950.0, b, None, st) == -1
        # the largest grid it holds still works
        grid = (C.c_uint * (1024 * 1024))()
        assert f(ctx, 1, 0, 1024, 1024, 0, 1, 950.0, b, grid, st) == 1
        assert np.frombuffer(grid, dtype=np.uint32).sum() == int(st[11]) > 0
    finally:
        hip_lib.st_free_context(ctx)
