"""The drop-in boundary as the reference's callers drive it, on the GPU:

* pysoltrace's calling convention (app/deploy/api/pysoltrace.py): no
  argtypes, restype only; `char` arguments passed as c_wchar (:430, 527,
  769-771), st_uint_t indices as c_uint32 (:428, 746, 765-776), a 3-element
  grating array where st_optic reads 4 (:406 vs coretrace/stapi.cpp:190-191),
  the seed as c_uint16 (:1599) and a callback declared with c_uint32
  arguments (:15).  The run must equal the same scene set through the
  declared prototypes;
* cancellation: a callback returning 0 stops the trace, and st_sim_run
  returns the records made so far (coretrace/raytrace.cpp:645-648,
  stapi.cpp:709-727) -- a consistent prefix, not an error;
* the GUI's threading (app/src/trace.cpp:709-730): one context per host
  thread, N/T rays each with the remainder on the first, seeds growing by
  123 i; every context traced concurrently equals its own run alone."""
import ctypes as C
import threading

import numpy as np
import pytest

from soltrace_amd import stapi, workloads
from harness import run_scene

pytestmark = pytest.mark.gpu

CB32 = C.CFUNCTYPE(C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p)


def _apply_pysoltrace_style(pdll, p, sc):
    """Set scene `sc` through the C API the way pysoltrace's Optics.Create,
    Sun.Create, Stage.Create and Element.Create call it."""
    for fn in ("st_add_optic", "st_optic", "st_sun", "st_sun_xyz", "st_sun_userdata", "st_add_stage",
               "st_stage_xyz", "st_stage_aim", "st_stage_zrot", "st_stage_flags", "st_add_element",
               "st_element_enabled", "st_element_xyz", "st_element_aim", "st_element_zrot", "st_element_aperture",
               "st_element_aperture_params", "st_element_surface", "st_element_surface_params",
               "st_element_interaction", "st_element_optic", "st_sim_errors", "st_sim_params"):
        getattr(pdll, fn).restype = C.c_int
    pv = C.c_void_p(p)
    for oid, o in enumerate(sc.optics):
        pdll.st_add_optic(pv, C.c_char_p(o.name.encode()))
        grating = (C.c_double * 3)()
        for i, f in enumerate((o.front, o.back)):
            ra = (C.c_double * len(f.refl_table or []))()
            rv = (C.c_double * len(f.refl_table or []))()
            ta = (C.c_double * len(f.trans_table or []))()
            tv = (C.c_double * len(f.trans_table or []))()
            rc = pdll.st_optic(pv, C.c_uint32(oid), C.c_int(i + 1), C.c_wchar(f.dist[0]), C.c_int(1), C.c_int(3),
                               C.c_int(4), C.c_double(f.rreal), C.c_double(0.0), C.c_double(f.refl),
                               C.c_double(f.trans), C.pointer(grating), C.c_double(f.slope), C.c_double(f.spec),
                               C.c_int(0), C.c_int(0), C.pointer(ra), C.pointer(rv), C.c_int(0), C.c_int(0),
                               C.pointer(ta), C.pointer(tv))
            assert rc == 1
    assert pdll.st_sun(pv, C.c_int(int(sc.point_source)), C.c_wchar(sc.sun_shape[0]), C.c_double(sc.sun_sigma)) == 1
    assert pdll.st_sun_xyz(pv, *[C.c_double(v) for v in sc.sun_xyz]) == 1
    if sc.sun_table is not None:
        a = (C.c_double * len(sc.sun_table[0]))(*sc.sun_table[0])
        v = (C.c_double * len(sc.sun_table[1]))(*sc.sun_table[1])
        assert pdll.st_sun_userdata(pv, C.c_uint32(len(sc.sun_table[0])), C.pointer(a), C.pointer(v)) == len(a)
    for si, st in enumerate(sc.stages):
        pdll.st_add_stage(pv)
        sid = C.c_uint32(si)
        pdll.st_stage_xyz(pv, sid, *[C.c_double(v) for v in st.xyz])
        pdll.st_stage_aim(pv, sid, *[C.c_double(v) for v in st.aim])
        pdll.st_stage_zrot(pv, sid, C.c_double(st.zrot))
        pdll.st_stage_flags(pv, sid, C.c_int(int(st.virtual)), C.c_int(int(st.multihit)),
                            C.c_int(int(st.tracethrough)))
        for ei, el in enumerate(st.elements):
            pdll.st_add_element(pv, sid)
            eid = C.c_uint32(ei)
            ap = (C.c_double * 8)(*(list(el.ap_params) + [0.0] * (8 - len(el.ap_params))))
            sp = (C.c_double * 8)(*(list(el.surf_params) + [0.0] * (8 - len(el.surf_params))))
            assert pdll.st_element_enabled(pv, sid, eid, C.c_int(int(el.enabled))) == 1
            assert pdll.st_element_xyz(pv, sid, eid, *[C.c_double(v) for v in el.xyz]) == 1
            assert pdll.st_element_aim(pv, sid, eid, *[C.c_double(v) for v in el.aim]) == 1
            assert pdll.st_element_zrot(pv, sid, eid, C.c_double(el.zrot)) == 1
            assert pdll.st_element_aperture(pv, sid, eid, C.c_wchar(el.aperture[0])) == 1
            assert pdll.st_element_aperture_params(pv, sid, eid, C.pointer(ap)) == 1
            assert pdll.st_element_surface(pv, sid, eid, C.c_wchar(el.surface[0])) == 1
            assert pdll.st_element_surface_params(pv, sid, eid, C.pointer(sp)) == 1
            assert pdll.st_element_interaction(pv, sid, eid, C.c_int(el.interaction)) == 1
            assert pdll.st_element_optic(pv, sid, eid, C.c_char_p(el.optic.encode())) == 1
    pdll.st_sim_errors(pv, C.c_int(int(sc.sunshape_err)), C.c_int(int(sc.optics_err)))
    pdll.st_sim_params(pv, C.c_int(int(sc.raycount)), C.c_int(int(sc.maxcount)), C.c_int(int(sc.dyn_group)))


def _records_pysoltrace_style(pdll, p):
    pv = C.c_void_p(p)
    pdll.st_num_intersections.restype = C.c_int
    n = pdll.st_num_intersections(pv)
    cols = {k: (C.c_double * n)() for k in ("x", "y", "z", "cx", "cy", "cz")}
    ints = {k: (C.c_int * n)() for k in ("element", "stage", "number")}
    for fn in ("st_locations", "st_cosines", "st_elementmap", "st_stagemap", "st_raynumbers", "st_sun_stats"):
        getattr(pdll, fn).restype = C.c_int
    pdll.st_locations(pv, C.pointer(cols["x"]), C.pointer(cols["y"]), C.pointer(cols["z"]))
    pdll.st_cosines(pv, C.pointer(cols["cx"]), C.pointer(cols["cy"]), C.pointer(cols["cz"]))
    pdll.st_elementmap(pv, C.pointer(ints["element"]))
    pdll.st_stagemap(pv, C.pointer(ints["stage"]))
    pdll.st_raynumbers(pv, C.pointer(ints["number"]))
    b = [C.c_double() for _ in range(4)]
    ns = C.c_int()
    pdll.st_sun_stats(pv, *[C.pointer(v) for v in b], C.pointer(ns))
    pos = np.stack([np.frombuffer(cols[k], dtype=np.float64) for k in ("x", "y", "z")], axis=1)
    cos = np.stack([np.frombuffer(cols[k], dtype=np.float64) for k in ("cx", "cy", "cz")], axis=1)
    el, stg, num = (np.frombuffer(ints[k], dtype=np.int32).copy() for k in ("element", "stage", "number"))
    return pos, cos, el, stg, num, [v.value for v in b], ns.value


def test_pysoltrace_calling_convention(hip_lib):
    sc = workloads.tower(n_helio=50, raycount=20000, r_max=400.0)
    seed = 54321  # as c_uint16
    pdll = C.CDLL(stapi.DEFAULT_LIB)  # its own handle: no argtypes declared
    pdll.st_create_context.restype = C.c_void_p
    p = pdll.st_create_context()
    calls = []

    @CB32
    def cb(ntracedtotal, ntraced, ntotrace, curstage, nstages, data):
        calls.append((ntraced, ntotrace, curstage, nstages))
        return 1

    try:
        _apply_pysoltrace_style(pdll, p, sc)
        pdll.st_sim_run.restype = C.c_int
        n = pdll.st_sim_run(C.c_void_p(p), C.c_uint16(seed), cb, 0)
        assert n > 20000
        pos, cos, el, stg, num, box, nsun = _records_pysoltrace_style(pdll, p)
    finally:
        pdll.st_free_context.restype = C.c_bool
        pdll.st_free_context(C.c_void_p(p))
    assert calls and calls[0][1] == 20000 and all(c[3] == 2 for c in calls)
    n_r, r, msgs, _ = run_scene(hip_lib, sc, seed)
    assert n_r == n, msgs
    np.testing.assert_array_equal(pos, r.pos)
    np.testing.assert_array_equal(cos, r.cos)
    np.testing.assert_array_equal(el, r.element)
    np.testing.assert_array_equal(stg, r.stage)
    np.testing.assert_array_equal(num, r.raynum)
    assert nsun == r.nsunrays and box == list(r.sun_box)


@pytest.mark.parametrize("bits", [64, 32])
def test_cancel_returns_the_partial_records(hip_lib, bits):
    # 2e6 hits on the tower need several stage-1 batches; cancelling at the
    # first callback keeps the first batch's stage-1 records
    sc = workloads.tower(n_helio=200, raycount=2_000_000, r_max=600.0)
    seed = 7

    def setup(lib, ctx):
        lib.dll.st_amd_set_precision(ctx, bits)

    full = run_scene(hip_lib, sc, seed, setup=setup)
    ctx = hip_lib.st_create_context()
    try:
        sc.apply(hip_lib, ctx)
        setup(hip_lib, ctx)
        seen = []

        @stapi.CALLBACK
        def cancel(ntracedtotal, ntraced, ntotrace, curstage, nstages, data):
            seen.append((ntraced, ntotrace, curstage))
            return 0

        n = hip_lib.st_sim_run(ctx, seed, cancel, None)
        part = stapi.fetch_results(hip_lib, ctx)
    finally:
        hip_lib.st_free_context(ctx)
    assert len(seen) == 1 and seen[0][2] == 1 and 0 < seen[0][0] < 2_000_000
    assert 0 < n == len(part) < full[0], (n, full[0])
    assert (part.stage == 1).all()
    k = seen[0][0]  # rays traced before the cancel
    np.testing.assert_array_equal(np.unique(part.raynum), np.arange(1, k + 1))
    f = full[1]
    keep = (f.stage == 1) & (f.raynum <= k)
    np.testing.assert_array_equal(part.element, f.element[keep])
    np.testing.assert_array_equal(part.raynum, f.raynum[keep])
    np.testing.assert_array_equal(part.pos, f.pos[keep])


def test_contexts_on_host_threads_equal_their_runs_alone(hip_lib):
    # the GUI's RunTraceMultiThreaded scheme with 4 threads on one GPU
    T, N, seed0 = 4, 120_001, 123
    jobs, seed = [], seed0
    for i in range(T):
        rays = N // T + (N % T if i == 0 else 0)
        seed += i * 123
        jobs.append((rays, seed))
    results = [None] * T

    def work(i):
        rays, s = jobs[i]
        sc = workloads.tower(n_helio=200, raycount=rays, r_max=600.0)
        results[i] = run_scene(hip_lib, sc, s)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(T)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert [j[1] for j in jobs] == [123, 246, 492, 861]
    for i, (rays, s) in enumerate(jobs):
        sc = workloads.tower(n_helio=200, raycount=rays, r_max=600.0)
        alone = run_scene(hip_lib, sc, s)
        conc = results[i]
        assert conc is not None and conc[0] == alone[0] > rays, (i, conc and conc[2], alone[2])
        for f in ("element", "stage", "raynum"):
            np.testing.assert_array_equal(getattr(conc[1], f), getattr(alone[1], f))
        np.testing.assert_array_equal(conc[1].pos, alone[1].pos)
        assert conc[1].nsunrays == alone[1].nsunrays
