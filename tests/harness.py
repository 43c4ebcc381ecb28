"""Drive any stapi library (HIP build, oracle, reference build) with a Scene."""
import ctypes as C

import numpy as np

from soltrace_amd import stapi


def run_scene(lib, scene, seed, setup=None, after=None):
    ctx = lib.st_create_context()
    assert ctx, "st_create_context returned NULL"
    try:
        scene.apply(lib, ctx)
        if setup is not None:
            setup(lib, ctx)
        n = stapi.run(lib, ctx, seed)
        res = stapi.fetch_results(lib, ctx) if n >= 0 else None
        msgs = stapi.messages(lib, ctx)
        extra = after(lib, ctx) if after is not None else None
        return n, res, msgs, extra
    finally:
        lib.st_free_context(ctx)


def counter_mode(lib, ctx):
    """Oracle set-up for comparisons with the HIP library: the per-ray Philox
    streams, and torus solves started where the library starts them for
    origins far outside the torus."""
    for name in ("st_oracle_set_rng", "st_oracle_set_torus_start"):
        f = getattr(lib.dll, name)
        f.argtypes = [C.c_void_p, C.c_int]
        f.restype = C.c_int
    assert lib.dll.st_oracle_set_rng(ctx, 1) == 1
    assert lib.dll.st_oracle_set_torus_start(ctx, 1) == 1


def oracle_sun_index(lib, ctx):
    n = lib.st_num_intersections(ctx)
    out = (C.c_uint64 * max(1, n))()
    k = lib.dll.st_oracle_ray_sun_index(ctx, out)
    return np.frombuffer(out, dtype=np.uint64, count=k).copy()


def hip_sun_index(lib, ctx, n):
    out = (C.c_ulonglong * max(1, n))()
    k = lib.dll.st_amd_ray_sun_index(ctx, out, n)
    return np.frombuffer(out, dtype=np.uint64, count=max(k, 0)).copy()


def assert_records_match(a, b, pos_tol, cos_tol):
    assert len(a) == len(b), f"record counts differ: {len(a)} vs {len(b)}"
    np.testing.assert_array_equal(a.element, b.element)
    np.testing.assert_array_equal(a.stage, b.stage)
    np.testing.assert_array_equal(a.raynum, b.raynum)
    np.testing.assert_allclose(a.pos, b.pos, rtol=0, atol=pos_tol)
    np.testing.assert_allclose(a.cos, b.cos, rtol=0, atol=cos_tol)
    assert a.nsunrays == b.nsunrays
    np.testing.assert_allclose(a.sun_box, b.sun_box, rtol=1e-12, atol=1e-9)
