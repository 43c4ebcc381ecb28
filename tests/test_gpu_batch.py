"""st_amd_sim_run_batch (include/stapi_amd.h): several contexts' scenes
traced together by persistent launches that drain per-wave tiles of every
scene.  Each scene's records must be those of its own st_sim_run with the
same seed: bit for bit in FP64, to 1e-4 m with equal hit sequences in FP32.
The scenes are the batched workload's mix at reduced sizes (tower fields of
200-1,500 heliostats and sun-tracking dishes of toroidal sectors and order-6
Zernike facets; suns from st_sun_position at 35 N), of uneven ray counts and
two kernel variants.  One dish is also held to the oracle, pinning that
scene type itself; a scene that fails in the batch fails alone, with its own
message, as its own run would."""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

from soltrace_amd import stapi, workloads
from harness import assert_records_match, counter_mode, run_scene

pytestmark = pytest.mark.gpu


def _batch(lib, scenes, seeds, bits):
    ctxs = []
    try:
        for sc in scenes:
            c = lib.st_create_context()
            sc.apply(lib, c)
            lib.dll.st_amd_set_precision(c, bits)
            ctxs.append(c)
        rc, counts = stapi.run_batch(lib, ctxs, seeds)
        res = [stapi.fetch_results(lib, c) if n >= 0 else None for c, n in zip(ctxs, counts)]
        msgs = [stapi.messages(lib, c) for c in ctxs]
        return rc, counts, res, msgs
    finally:
        for c in ctxs:
            lib.st_free_context(c)


def _bitexact(a, b):
    assert len(a) == len(b)
    for f in ("pos", "cos"):
        assert np.array_equal(getattr(a, f).view(np.uint64), getattr(b, f).view(np.uint64)), f
    for f in ("element", "stage", "raynum"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    assert a.nsunrays == b.nsunrays and a.sun_box == b.sun_box


@pytest.mark.parametrize("bits", [64, 32])
def test_batch_equals_separate_runs(hip_lib, bits):
    with tempfile.TemporaryDirectory() as d:
        scenes = workloads.batch_scenes(24, seed=7, hits=(3000, 30000), mon_dir=d, helio=(200, 1500))
        seeds = [1000 + 17 * k for k in range(len(scenes))]
        print(f"batch of {len(scenes)} scenes, FP{bits}", flush=True)
        rc, counts, res, msgs = _batch(hip_lib, scenes, seeds, bits)
        assert rc == 1 and min(counts) > 0, (counts, [m[-1:] for m in msgs])
        kinds = set()
        for k, (sc, sd) in enumerate(zip(scenes, seeds)):
            kinds.add(sc.stages[0].elements[0].surface)
            n, alone, m, _ = run_scene(hip_lib, sc, sd, setup=lambda lib, c: lib.dll.st_amd_set_precision(c, bits))
            assert n == counts[k], (k, n, counts[k], m)
            if bits == 64:
                _bitexact(res[k], alone)
            else:
                assert_records_match(res[k], alone, pos_tol=1e-4, cos_tol=1e-4)
        assert kinds == {"s", "m"}  # fields (spheres) and dishes (Zernike facets, torus sections)
    out = (C.c_int * 48)()
    ntuned = hip_lib.dll.st_amd_batch_tuning(out, 48)
    assert ntuned >= 2, ntuned
    for i in range(min(ntuned, 16)):
        key, tile, waves = out[3 * i], out[3 * i + 1], out[3 * i + 2]
        assert key // 100 in (32, 64) and tile in (64, 128, 256, 512) and waves in (1, 2, 3, 4)


def test_dish_equals_oracle(hip_lib, oracle_lib):
    with tempfile.TemporaryDirectory() as d:
        rng = np.random.default_rng(11)
        sc = workloads.dish(workloads.sun_position(35.0, 120.0, 10.5), 4000, d, rng)
        n_o, r_o, _, _ = run_scene(oracle_lib, sc, 5, setup=counter_mode)
        n_g, r_g, m_g, _ = run_scene(hip_lib, sc, 5, setup=lambda lib, c: lib.dll.st_amd_set_precision(c, 64))
        assert n_g == n_o > 4000, (n_g, n_o, m_g)
        assert_records_match(r_g, r_o, pos_tol=1e-6, cos_tol=1e-6)
        zer = (np.abs(r_o.element) <= 7) & (r_o.stage == 1)
        tor = (np.abs(r_o.element) >= 8) & (r_o.stage == 1)
        assert zer.sum() > 2000 and tor.sum() > 100, (zer.sum(), tor.sum())
        assert (r_o.element[r_o.stage == 2] == -1).sum() > 2000  # the receiver absorbs most of the light


def test_a_failing_scene_fails_alone(hip_lib):
    with tempfile.TemporaryDirectory() as d:
        scenes = workloads.batch_scenes(6, seed=3, hits=(2000, 6000), mon_dir=d, helio=(200, 400))
        scenes[2].maxcount = 100  # sun-ray cap reached (coretrace/raytrace.cpp:600-604)
        seeds = list(range(6))
        rc, counts, res, msgs = _batch(hip_lib, scenes, seeds, 64)
        assert rc == -1 and counts[2] == -1 and "maximum count" in msgs[2][-1]
        for k in (0, 1, 3, 4, 5):
            n, alone, _, _ = run_scene(hip_lib, scenes[k], seeds[k])
            assert counts[k] == n > 0
            _bitexact(res[k], alone)
