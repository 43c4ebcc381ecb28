"""Stage-data save/resume (the reference's st_sim_run_data,
coretrace/stapi.cpp:680-728; raytrace.cpp:544-578, 1044-1078;
generateray.cpp:165-219), oracle against the reference build, bit for bit.
The reference's C++ entry point is reached through oracle/_ref/ref_run_data.so
(plain arrays in, vectors to the reference).  The scene has three stages so
that the saved stage-2 array holds both kinds of row the reference leaves in
it: rays written out of stage 2 and rays read into it."""
import ctypes as C
import os

import numpy as np
import pytest

from soltrace_amd import stapi
from soltrace_amd.scene import Element, Optic, OpticFace, Scene, Stage

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SHIM = os.path.join(ROOT, "oracle", "_ref", "ref_run_data.so")


def three_stage_scene(n=1500, flat=False, n_stages=3):
    """Twelve parabolic (or flat) facets, a partly reflecting secondary above
    them, and an absorbing receiver between (n_stages 2: no receiver, so
    stage 2 is the last)."""
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=(0.1, -0.2, 1.0), raycount=n, maxcount=10 ** 6)
    s.optics = [Optic("mirror", OpticFace(refl=0.9, slope=0.8, spec=0.3), OpticFace(refl=0.9, slope=0.8, spec=0.3)),
                Optic("second", OpticFace(refl=0.6, slope=0.5), OpticFace(refl=0.6, slope=0.5)),
                Optic("recv", OpticFace(refl=0.1), OpticFace(refl=0.1))]
    facets = [Element(xyz=(x, y, 0.0), aim=(-0.02 * x, -0.02 * y, 1.0), aperture="r", ap_params=(1.0, 1.0),
                      surface="f" if flat else "p", surf_params=() if flat else (0.12, 0.12), optic="mirror")
              for x in (-1.65, -0.55, 0.55, 1.65) for y in (-1.1, 0.0, 1.1)]
    second = Element(xyz=(0, 0, 4.0), aim=(0.05, 0, -1.0), aperture="c", ap_params=(2.2,), surface="f",
                     optic="second")
    recv = Element(xyz=(0, 0, 1.5), aim=(0, 0, 1.0), aperture="r", ap_params=(1.6, 1.6), surface="f", optic="recv")
    s.stages = [Stage(elements=facets), Stage(elements=[second]), Stage(elements=[recv])][:n_stages]
    return s


def run_data(lib, scene, seed, s1=None, s2=None, save=False, api=None, setup=None):
    """(record count, results, saved rows, messages) of one st_sim_run_data."""
    ctx = lib.st_create_context()
    try:
        scene.apply(lib, ctx)
        if setup is not None:
            setup(lib, ctx)
        n = stapi.run_data(lib, ctx, seed, s1, s2, save, api=api)
        res = stapi.fetch_results(lib, ctx) if n >= 0 else None
        saved = stapi.saved_stage_data(lib, ctx, api=api) if n >= 0 else None
        return n, res, saved, stapi.messages(lib, ctx)
    finally:
        lib.st_free_context(ctx)


@pytest.fixture(scope="module")
def ref_api(ref_lib):
    if not os.path.exists(REF_SHIM):
        pytest.skip("make -C oracle ref")
    return C.CDLL(REF_SHIM)


def bitexact(a, b):
    assert a[0] == b[0] and a[0] >= 0, (a[0], b[0], a[3], b[3])
    ra, rb = a[1], b[1]
    for f in ("pos", "cos"):
        assert np.array_equal(getattr(ra, f).view(np.uint64), getattr(rb, f).view(np.uint64)), f
    for f in ("element", "stage", "raynum"):
        assert np.array_equal(getattr(ra, f), getattr(rb, f)), f
    assert ra.nsunrays == rb.nsunrays and ra.sun_box == rb.sun_box
    for x, y in zip(a[2], b[2]):
        assert x.shape == y.shape and np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_save_with_stage2_last_equals_reference(ref_lib, ref_api, oracle_lib):
    # stage 2 is the last stage: its rays go nowhere, but the saved stage-2
    # array still holds them (coretrace/raytrace.cpp:1065-1078)
    sc = three_stage_scene(n_stages=2)
    r = run_data(ref_lib, sc, 3, save=True, api=ref_api)
    o = run_data(oracle_lib, sc, 3, save=True)
    bitexact(o, r)
    assert len(o[2][1]) > 0


@pytest.mark.parametrize("seed", [3, 11])
def test_save_equals_reference(ref_lib, ref_api, oracle_lib, seed):
    sc = three_stage_scene()
    r = run_data(ref_lib, sc, seed, save=True, api=ref_api)
    o = run_data(oracle_lib, sc, seed, save=True)
    bitexact(o, r)
    s1, s2 = o[2]
    # stage-1 rows are the stage-1 records; the stage-2 array is as long as
    # the rays that entered stage 2, and holds some rays written out of it
    # (they carry stage-3 ray numbers) ahead of rays read into it
    res = o[1]
    st1 = res.stage == 1
    assert len(s1) == st1.sum() and len(s2) > 0
    assert set(s1[:, 7].astype(int)) == set(res.raynum[st1])
    into3 = set(res.raynum[res.stage == 3])
    assert into3 and into3 <= set(s2[:, 6].astype(int))


def test_load_equals_reference(ref_lib, ref_api, oracle_lib):
    sc = three_stage_scene()
    _, _, (s1, s2), _ = run_data(oracle_lib, sc, 5, save=True)
    r = run_data(ref_lib, sc, 9, s1, s2, api=ref_api)
    o = run_data(oracle_lib, sc, 9, s1, s2)
    bitexact(o, r)
    res = o[1]
    # stage 1 is the loaded rows, and the sun-ray count the last one's ray number
    st1 = res.stage == 1
    assert st1.sum() == len(s1) and res.nsunrays == int(s1[-1, 7])
    assert np.array_equal(np.sort(res.element[st1]), np.sort(s1[:, 6].astype(np.int32)))


def test_load_and_save_equals_reference(ref_lib, ref_api, oracle_lib):
    # both at once: the reference saves the loaded stage-1 rows again
    sc = three_stage_scene(800)
    _, _, (s1, s2), _ = run_data(oracle_lib, sc, 2, save=True)
    r = run_data(ref_lib, sc, 4, s1, s2, save=True, api=ref_api)
    o = run_data(oracle_lib, sc, 4, s1, s2, save=True)
    bitexact(o, r)
    assert np.array_equal(o[2][0], s1)


def test_stage2_rows_alone_set_the_ray_count(ref_lib, ref_api, oracle_lib):
    # without stage-1 rows nothing is loaded, but the ray count becomes the
    # number of stage-2 rows (coretrace/stapi.cpp:697-700)
    sc = three_stage_scene(5000)
    s2 = np.zeros((321, 7))
    r = run_data(ref_lib, sc, 8, None, s2, api=ref_api)
    o = run_data(oracle_lib, sc, 8, None, s2)
    bitexact(o, r)
    res = o[1]
    assert res.raynum[res.stage == 1].max() == 321


def test_zero_last_ray_number_fails_as_reference(ref_lib, ref_api, oracle_lib):
    sc = three_stage_scene(500)
    _, _, (s1, s2), _ = run_data(oracle_lib, sc, 5, save=True)
    s2 = s2.copy()
    s2[-1, 6] = 0.0
    r = run_data(ref_lib, sc, 9, s1, s2, api=ref_api)
    o = run_data(oracle_lib, sc, 9, s1, s2)
    assert r[0] == o[0] == -1
    assert r[3][-1] == o[3][-1] and o[3][-1].startswith("LastRayNumberInPreviousStage=0, stage 1")
