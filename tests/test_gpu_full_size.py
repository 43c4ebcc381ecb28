"""The three single-GPU workloads at their full sizes, checked through
properties that do not depend on the oracle finishing:
  * trough: 1 parabolic trough and its receiver tube, 1e6 ray hits;
  * tower: 1,000 spherical heliostats and a flat receiver, 1e7 ray hits;
  * faceted field: 10,000 heliostats of 4 flat facets (40,000 elements) and a
    cylindrical receiver, 5e7 ray hits.
Each runs in FP64 and FP32 with the same seed, and must satisfy:
  * every ray number 1..N has a stage-1 record (N hitting sun rays,
    coretrace/raytrace.cpp:780-797) and no stage-1 record has another;
  * the sun-ray count is at least N and the sun box is not empty;
  * the receiver flux maps of the two precisions agree to an NRMSE (RMS
    difference over the FP64 peak) of at most 0.5%, and their intercepted
    counts within 3 sigma of a binomial count -- the two paths draw the same
    errors, so both figures are far inside the bounds."""
import ctypes as C

import numpy as np
import pytest

from soltrace_amd import stapi, workloads

pytestmark = pytest.mark.gpu

NRMSE_MAX = 0.005


def _prepare(lib):
    d = lib.dll
    d.st_amd_set_precision.argtypes = [C.c_void_p, C.c_int]
    d.st_amd_flux_request.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double]
    d.st_amd_flux_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_double)]


def _run(lib, scene, bits, nx, ny, half_x, half_y, seed=2024):
    """Records' stage and ray-number maps, sun statistics and the receiver
    flux map (absorbed last records on stage 2, element 1) of one run."""
    d = lib.dll
    ctx = lib.st_create_context()
    try:
        scene.apply(lib, ctx)
        assert d.st_amd_set_precision(ctx, bits) == 1
        assert d.st_amd_flux_request(ctx, 1, 0, nx, ny, 1, -half_x, half_x, -half_y, half_y) == 1
        n = stapi.run(lib, ctx, seed)
        msgs = stapi.messages(lib, ctx)
        assert n > 0, msgs
        stage = np.zeros(n, dtype=np.int32)
        raynum = np.zeros(n, dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        assert lib.st_stagemap(ctx, ip(stage)) == 1 and lib.st_raynumbers(ctx, ip(raynum)) == 1
        b = [C.c_double() for _ in range(4)]
        ns = C.c_int()
        lib.st_sun_stats(ctx, *[C.byref(v) for v in b], C.byref(ns))
        flux = (C.c_uint * (nx * ny))()
        assert d.st_amd_flux_result(ctx, flux, None) == nx * ny
        return stage, raynum, [v.value for v in b], ns.value, np.frombuffer(flux, dtype=np.uint32).astype(float)
    finally:
        lib.st_free_context(ctx)


def _check(hip_lib, name, scene, nx, ny, half_x, half_y):
    _prepare(hip_lib)
    N = scene.raycount
    maps = {}
    for bits in (64, 32):
        print(f"{name}: {N:.0e} hits, FP{bits}", flush=True)
        stage, raynum, box, nsun, flux = _run(hip_lib, scene, bits, nx, ny, half_x, half_y)
        s1 = raynum[stage == 1]
        cover = np.bincount(s1, minlength=N + 1)
        assert s1.min() >= 1 and s1.max() == N, (s1.min(), s1.max())
        assert np.all(cover[1:] > 0), int(np.sum(cover[1:] == 0))
        assert nsun >= N and box[0] < box[1] and box[2] < box[3], (nsun, box)
        maps[bits] = flux
        del stage, raynum
    f64, f32 = maps[64], maps[32]
    assert f64.sum() > 0
    nrmse = float(np.sqrt(np.mean((f32 - f64) ** 2)) / f64.max())
    c64, c32 = f64.sum(), f32.sum()
    sigma = np.sqrt(c64 * max(1e-12, 1.0 - c64 / N))
    print(f"{name}: flux NRMSE {nrmse:.2e}, intercepted {c64:.0f} / {c32:.0f}", flush=True)
    assert nrmse <= NRMSE_MAX
    assert abs(c32 - c64) <= 3.0 * np.sqrt(2.0) * sigma


def test_trough_full_size(hip_lib):
    _check(hip_lib, "trough", workloads.trough(raycount=1_000_000), 50, 50, 0.5, 6.0)


def test_tower_full_size(hip_lib):
    _check(hip_lib, "tower", workloads.tower(n_helio=1000, raycount=10_000_000), 100, 100, 7.5, 7.5)


def test_faceted_field_full_size(hip_lib):
    sc = workloads.faceted_field(n_helio=10_000, raycount=50_000_000)
    assert sum(len(s.elements) for s in sc.stages[:1]) == 40_000
    _check(hip_lib, "faceted field", sc, 128, 256, 30.0, 10.0)
