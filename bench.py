"""Benchmark: ray hits/s of the coretrace trace loop on one MI355X.

Workload (one "step" = one st_sim_run of the whole job): the tower heliostat
field -- 1,000 spherically focused 10 x 10 m heliostats on a radially
staggered layout within 1 km, each aimed at a flat 15 x 15 m receiver at
150 m; Buie CSR 0.1 sun (tabulated 'd' profile), 1.5 mrad slope error, sun
at 45 deg elevation / 180 deg azimuth; 1e7 ray hits; 100 x 100 absorbed-flux
map on the receiver binned on the GPU.  Inputs are synthetic and generated by
soltrace_amd.workloads.tower.

The headline value is the FP32-bulk path.  The same job is also run on the
FP64 path with the same seed (same counter-based draws) to report the FP32 /
FP64 speed ratio and the flux-map NRMSE = RMS(F32 - F64) / max(F64).  The
other two single-GPU workloads run beside it ("workloads"): the parabolic
trough (1e6 hits, FP64, the reference's plain path) and the 40,000-facet
field (5e7 hits, FP32).  "end_to_end" adds the result getters' copy of every
record to the host (st_locations .. st_raynumbers), which st_sim_run itself
leaves in HBM.

Roofline: the dominant kernel's achieved rate is its ALGORITHMIC floating-
point work per launch over its average launch time (HIP events on its own
stream, summed over the timed steps).  The work per unit comes from the
oracle built with operation counting (oracle/_build/st_oracle_count.so, run
on a bounded sample of the same workload on the per-ray streams): flops per
sun ray, per hitting ray, per ray entering a later stage, per tabulated-
sunshape draw, and per element test of each stage; the element tests a
kernel makes are the GPU's own event counts (its cull is exact, so these are
the tests the search needs); BVH node visits are counted as no flops.

"batched" traces 256 uneven scenes (tower fields of 200-5,000 heliostats and
sun-tracking dishes with Zernike facets and torus sections, suns over a year
at 35 N, 1e5-2e6 ray hits each) in one st_amd_sim_run_batch call per step,
beside the same contexts traced one st_sim_run at a time.

"surrogate" is the learned flux surrogate on the same field: 200,000
per-heliostat receiver images traced by the library train a BF16 MLP on the
GPU's MFMA units, which then drives a 500-step aim-point sweep with the
tracer as fallback and judge; its speed per step is compared with the same
sweep scoring each candidate by tracing, and its final peak flux with a
1e8-hit FP64 trace (soltrace_amd/surrogate.py).

`cpu_baseline` times the reference's own coretrace, compiled from source in
this repository's oracle/ recipe (oracle/_ref/coretrace_ref.so), on the
host's cores with the GUI's multithreaded scheme (one context per thread,
rays split evenly, seeds seed, seed+123, seed+369, ...), on a bounded sample
of the same scene, and on one thread.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-workloads] [--no-batch] [--no-surrogate]
       [--only batched|surrogate|mlp]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from soltrace_amd import stapi, workloads  # noqa: E402

METRIC = "ray hits/sec on 1 MI355X (FP32 vs FP64 vs CPU); flux-map NRMSE vs FP64"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (published; needs packed FP32)
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector peak; also the non-packed FP32 rate
N_HITS = 10_000_000
NX = NY = 100
COUNT_SO = os.path.join(ROOT, "oracle", "_build", "st_oracle_count.so")
# CPU threads the box gives one GPU's job
BOX_CPU_SHARE = 16

# HBM bytes per unit (ray) of the dominant kernels on this workload, from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script, scaled
# by what the counters read for a plain 1 GiB device copy (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0); profiles/hbm_traffic_tower1000.md has the runs.  Not
# measured in this run: a PMC run cannot time the kernels.
PMC_BYTES_PER_UNIT = {
    "trace_stage0": 349.0 + 221.5,
    "sun_mask": 0.14 + 0.125,
}
PMC_SOURCE = "profiles/hbm_traffic_tower1000.md (calibrated FETCH_SIZE + WRITE_SIZE of an earlier run of this bench)"


def _lib():
    lib = stapi.load()
    d = lib.dll
    d.st_amd_set_precision.argtypes = [C.c_void_p, C.c_int]
    d.st_amd_flux_request.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double, C.c_double]
    d.st_amd_flux_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_double)]
    d.st_amd_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
    d.st_amd_events.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
    d.st_amd_count_events.argtypes = [C.c_void_p, C.c_int]
    return lib


def _context(lib, scene, bits):
    ctx = lib.st_create_context()
    if not ctx:
        raise RuntimeError("st_create_context failed (no GPU?)")
    scene.apply(lib, ctx)
    lib.dll.st_amd_set_precision(ctx, bits)
    half = 7.5
    if lib.dll.st_amd_flux_request(ctx, 1, 0, NX, NY, 1, -half, half, -half, half) != 1:
        raise RuntimeError("flux request failed")
    return ctx


def _run(lib, ctx, seed):
    n = stapi.run(lib, ctx, seed)
    if n < 0:
        raise RuntimeError("st_sim_run failed: " + "; ".join(stapi.messages(lib, ctx)))
    return n


def _flux(lib, ctx):
    buf = (C.c_uint * (NX * NY))()
    lib.dll.st_amd_flux_result(ctx, buf, None)
    return np.frombuffer(buf, dtype=np.uint32).astype(np.float64).reshape(NX, NY)


def _arr(lib, fn, ctx, n):
    buf = (C.c_double * n)()
    getattr(lib.dll, fn)(ctx, buf, n)
    return list(buf)


def _timed_steps(lib, ctx, steps, warmup, seed, sync):
    """Mean wall time of `steps` runs, and the mean of each run's
    st_amd_timing (kernel times from HIP events on each kernel's stream)."""
    for _ in range(warmup):
        _run(lib, ctx, seed)
    sync()
    acc = np.zeros(53)
    t0 = time.perf_counter()
    for _ in range(steps):
        _run(lib, ctx, seed)
        acc += np.array(_arr(lib, "st_amd_timing", ctx, 53))
    sync()
    return (time.perf_counter() - t0) / steps, list(acc / steps)


def _workload_rate(lib, scene, bits, steps, seed, sync):
    """ray hits/s and ms per run of one workload (no flux map)."""
    ctx = lib.st_create_context()
    try:
        scene.apply(lib, ctx)
        lib.dll.st_amd_set_precision(ctx, bits)
        t, tm = _timed_steps(lib, ctx, steps, 1, seed, sync)
        return {"value": scene.raycount / t, "ms_per_step": t * 1e3, "precision": f"fp{bits}",
                "sun_rays": int(tm[19]), "records": int(sum(tm[36:52]))}
    finally:
        lib.st_free_context(ctx)


def _end_to_end(lib, ctx, seed, sync):
    """One run plus the result getters' copy of every record into arrays
    the caller allocates (st_locations, st_cosines, st_elementmap,
    st_stagemap, st_raynumbers, coretrace/stapi.cpp:564-659)."""
    sync()
    t0 = time.perf_counter()
    n = _run(lib, ctx, seed)
    t1 = time.perf_counter()
    d = [np.empty(n) for _ in range(6)]
    m = [np.empty(n, dtype=np.int32) for _ in range(3)]
    dp = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in d]
    ip = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in m]
    ok = (lib.st_locations(ctx, *dp[:3]) == 1 and lib.st_cosines(ctx, *dp[3:]) == 1 and
          lib.st_elementmap(ctx, ip[0]) == 1 and lib.st_stagemap(ctx, ip[1]) == 1 and
          lib.st_raynumbers(ctx, ip[2]) == 1)
    t2 = time.perf_counter()
    if not ok:
        raise RuntimeError("result getters failed")
    gb = n * (6 * 8 + 3 * 4) / 1e9
    return {"ms_per_step": (t2 - t0) * 1e3, "trace_ms": (t1 - t0) * 1e3, "getters_ms": (t2 - t1) * 1e3,
            "records": n, "getter_gb": gb, "getter_gb_s": gb / (t2 - t1), "value": N_HITS / (t2 - t0)}


def counted_work(scene_fn, hits, seed):
    """Algorithmic flops per unit from the operation-counting oracle on a
    bounded sample of the workload (per-ray streams, the GPU's draws).
    Returns per-unit figures: flops (add/sub, mul, div, sqrt) and
    transcendentals, or None when the counting oracle is not built."""
    if not os.path.exists(COUNT_SO):
        return None
    lib = stapi.StLib(COUNT_SO)
    for n in ("st_oracle_set_rng", "st_oracle_set_torus_start"):
        getattr(lib.dll, n).argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_opcount_read.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
    sc = scene_fn(hits)
    ctx = lib.st_create_context()
    try:
        sc.apply(lib, ctx)
        lib.dll.st_oracle_set_rng(ctx, 1)
        lib.dll.st_oracle_set_torus_start(ctx, 1)
        lib.dll.st_opcount_reset()
        t0 = time.perf_counter()
        if stapi.run(lib, ctx, seed) < 0:
            return None
        dt = time.perf_counter() - t0
        res = stapi.fetch_results(lib, ctx)
    finally:
        lib.st_free_context(ctx)
    out = (C.c_ulonglong * 83)()
    lib.dll.st_opcount_read(out, 83)
    a = np.frombuffer(out, dtype=np.uint64).astype(np.float64)
    ph, tests = a[:80].reshape(16, 5), a[80:]
    fl = ph[:, :4].sum(axis=1)
    tr = ph[:, 4]
    sun = float(res.nsunrays)
    rays2 = float(len(np.unique(res.raynum[res.stage == 2]))) if (res.stage == 2).any() else 0.0
    def per(p, units):
        return {"flops": fl[p] / units if units else 0.0, "trans": tr[p] / units if units else 0.0}
    return {
        "sample": f"{hits} ray hits ({int(sun)} sun rays) on the counting oracle, {dt:.1f} s",
        "sun_ray": {"flops": (fl[0] + fl[1]) / sun, "trans": (tr[0] + tr[1]) / sun},
        "hit": {"flops": (fl[1] + fl[2]) / hits, "trans": (tr[1] + tr[2]) / hits},
        "sunshape": per(3, hits),
        "stage2_ray": per(5, rays2),
        "test_stage1": per(12, tests[0]),
        "test_stage2": per(13, tests[1]),
    }


def kernel_work(cw, ev, hits_per_run, tm):
    """Algorithmic flops of each kernel in one run: the counted per-unit
    figures times the GPU's units and element tests (st_amd_events)."""
    base = 17
    mask = ev[base:base + 5]
    st = [ev[base + 5 * (1 + s): base + 5 * (2 + s)] for s in range(2)]
    out = {
        "sun_mask": {"launches": ev[0], "units": mask[0],
                     "flops": mask[0] * cw["sun_ray"]["flops"] + mask[1] * cw["test_stage1"]["flops"]},
        "trace_stage0": {"launches": ev[1], "units": st[0][0],
                         "flops": st[0][0] * cw["hit"]["flops"] + st[0][1] * cw["test_stage1"]["flops"]},
        "trace_stage1": {"launches": ev[2], "units": st[1][0],
                         "flops": st[1][0] * cw["stage2_ray"]["flops"] + st[1][1] * cw["test_stage2"]["flops"]},
        "sunshape": {"launches": ev[1], "units": hits_per_run, "flops": hits_per_run * cw["sunshape"]["flops"]},
    }
    for k, s_ in (("trace_stage0", 0), ("trace_stage1", 1)):
        out[k]["tests"] = st[s_][1]
        out[k]["bvh_nodes"] = st[s_][2]
    out["sun_mask"]["tests"] = mask[1]
    return out


def cpu_baseline(scene_fn, seed, target_s=12.0):
    """The reference build on the host cores, GUI threading scheme
    (app/src/trace.cpp:709-729), on a bounded sample of the same scene, and
    its single-thread rate."""
    so = os.path.join(ROOT, "oracle", "_ref", "coretrace_ref.so")
    kind = "reference"
    if not os.path.exists(so):
        so = os.path.join(ROOT, "oracle", "_build", "st_oracle.so")
        kind = "port"
    if not os.path.exists(so):
        return None
    ref = stapi.StLib(so)
    host_cpus = len(os.sched_getaffinity(0))
    threads = max(1, min(BOX_CPU_SHARE, host_cpus))
    # single thread, on a sample of about a second
    cal = 1500
    sc = scene_fn(cal)
    ctx = ref.st_create_context()
    sc.apply(ref, ctx)
    t0 = time.perf_counter()
    ref.st_sim_run(ctx, seed, stapi._NOCB, None)
    rate1 = cal / (time.perf_counter() - t0)
    ref.st_free_context(ctx)
    per_thread = int(max(1000, min(2_000_000, rate1 * target_s)))
    total = per_thread * threads
    ctxs, seeds = [], []
    sd = seed
    for i in range(threads):
        sd = sd + i * 123  # cumulative, as the GUI does
        seeds.append(sd)
        sc = scene_fn(per_thread)
        c = ref.st_create_context()
        sc.apply(ref, c)
        ctxs.append(c)
    res = [0] * threads

    def work(i):
        res[i] = ref.st_sim_run(ctxs[i], seeds[i] & 0xFFFFFFFF, stapi._NOCB, None)

    ths = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    for c in ctxs:
        ref.st_free_context(c)
    if min(res) < 0:
        return None
    return {"value": total / dt, "unit": "ray hits/s", "cores": threads, "kind": kind,
            "host_cpus": host_cpus, "value_1_thread": rate1,
            "sample": f"{threads} contexts x {per_thread} ray hits of the same 1,000-heliostat scene "
                      f"(dynamic grouping on) on {threads} threads ({host_cpus} CPUs visible, "
                      f"{BOX_CPU_SHARE} per GPU on the box), {dt:.1f} s wall; one thread: {rate1:.0f} hits/s "
                      f"on {cal} hits"}


def surrogate_workload(lib, seed, log):
    """The learned flux surrogate on the tower field: 200 random states
    (sun within 160-200 deg azimuth and 35-55 deg elevation, slope error
    1-2 mrad, aim offsets within +-6 m) x 1,000 heliostats = 200,000
    per-heliostat 64 x 64 receiver images at ~2e5 rays each (ten 2e7-hit
    traces per state, averaged), 3 held-out states at ~1e6 rays per image;
    the 8 -> 6 x 256 -> 4096 BF16 network trained by the library's MFMA
    kernels (20,000 Adam steps of 512); a 500-step aim-point sweep at the
    bench sun with the tracer as fallback (5e7-hit anchors) and judge; the
    same sweep scoring candidates by 1e7-hit traces, timed over 10 steps;
    start and end judged by 1e8-hit FP64 traces."""
    import torch
    from soltrace_amd import surrogate as S

    f = S.Field(lib, 1000, bits=32)
    t0 = time.perf_counter()
    X, T = S.make_dataset(f, 200, 20_000_000, seed, repeats=10, log=log)
    Xv, Tv = S.make_dataset(f, 3, 20_000_000, seed + 1, repeats=50)
    data_s = time.perf_counter() - t0
    sur = S.Surrogate(lib, max_batch=1024, seed=1)
    fit = sur.fit(X, T, 20_000, batch=512, val=(Xv, Tv), log=log)
    del X, T
    # the MFMA kernels' rate: training steps and forwards, HIP events on torch's stream (the one they launch on)
    n_in, width, depth, n_out = sur.net.shape
    macs = n_in * width + (depth - 1) * width * width + width * n_out
    xb = Xv[:512].contiguous()
    tb = Tv[:512].contiguous()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sur.net.train_step(xb, tb, 0.0)
    ev0.record()
    for _ in range(50):
        sur.net.train_step(xb, tb, 0.0)
    ev1.record()
    torch.cuda.synchronize()
    step_ms = ev0.elapsed_time(ev1) / 50
    xf = Xv[:1024].contiguous()
    sur.net.forward(xf)
    ev0.record()
    for _ in range(50):
        sur.net.forward(xf)
    ev1.record()
    torch.cuda.synchronize()
    fwd_ms = ev0.elapsed_time(ev1) / 50
    train_tflops = 3 * 2 * 512 * macs / (step_ms * 1e-3) / 1e12
    fwd_tflops = 2 * 1024 * macs / (fwd_ms * 1e-3) / 1e12

    st0 = S.FieldState(offsets=np.zeros((f.n, 2)))
    f.set_state(st0)
    r = S.aim_sweep(f, sur, st0.offsets, steps=500, anchor_hits=50_000_000, seed=seed)
    f.set_state(st0)
    rt = S.aim_sweep_traced(f, st0.offsets, steps=10, eval_hits=10_000_000, seed=seed)
    f64 = S.Field(lib, 1000, bits=64, by_source=False)
    j0 = float(S.judge(f64, st0.offsets, 100_000_000, seed + 7).max().item())
    j1 = float(S.judge(f64, r.offsets, 100_000_000, seed + 7).max().item())
    f.close()
    f64.close()
    sur.net.close()
    ms_net, ms_tr = r.wall_s / r.steps * 1e3, rt.wall_s / rt.steps * 1e3
    return {
        "workload": "tower field, 1,000 heliostats: 200,000 per-heliostat 64x64 receiver images at ~2e5 rays "
                    "each train an 8 -> 6x256 -> 4096 BF16 MLP (FP32 accumulate) on the GPU; 500-step "
                    "aim-point sweep with tracer fallback, judged by 1e8-hit FP64 traces",
        "images": 200 * 1000, "rays_per_image": 2e5, "data_s": data_s,
        "train_s": fit["train_s"], "train_steps": fit["steps"], "batch": fit["batch"],
        "validation": {k: fit[k] for k in ("rel_rms_bright", "rms_bright_wm2", "peak_rel_median", "peak_rel_p95",
                                           "power_rel_median", "power_rel_p95")},
        "mlp": {"train_step_ms": step_ms, "train_tflops": train_tflops, "forward_1024_ms": fwd_ms,
                "forward_tflops": fwd_tflops, "peak_tflops_bf16": 2500.0,
                "frac_train": train_tflops / 2500.0, "macs_per_row": macs},
        "sweep": {"steps": r.steps, "accepted": r.accepted, "fallbacks": r.fallbacks, "traces": r.traces,
                  "wall_s": r.wall_s, "ms_per_step": ms_net, "judged_errors_max": max(r.judged_errors or [0.0]),
                  "peak_pred_wm2": r.peak_pred, "peak_judged_start_wm2": j0, "peak_judged_end_wm2": j1,
                  "peak_error": abs(r.peak_pred - j1) / j1},
        "traced_sweep": {"steps": rt.steps, "traces": rt.traces, "ms_per_step": ms_tr, "eval_hits": 10_000_000},
        "speedup_per_step": ms_tr / ms_net,
        "targets": {"speedup": ">= 20", "peak_error": "<= 0.02"},
    }


def batch_workload(lib, seed, sync, log, n_scenes=256, bits=32, reps=2):
    """Batched uneven scenes: 256 scenes from a fixed seed, each a tower
    field of U[200, 5000] heliostats or a sun-tracking dish of order-6
    Zernike facets and torus sections, suns over a year at 35 N, U[1e5, 2e6]
    ray hits each (workloads.batch_scenes), traced by st_amd_sim_run_batch;
    one untimed batch (scene compile, launch-shape autotuning), then `reps`
    timed batches; the same contexts traced one by one with st_sim_run for
    comparison."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        scenes = workloads.batch_scenes(n_scenes, seed=2026, hits=(100_000, 2_000_000), mon_dir=d)
        ctxs = []
        try:
            for sc in scenes:
                c = lib.st_create_context()
                sc.apply(lib, c)
                lib.dll.st_amd_set_precision(c, bits)
                ctxs.append(c)
            setup_s = time.perf_counter() - t0
            log(f"batch: {len(ctxs)} scenes set up in {setup_s:.1f} s")
            seeds = [(seed + 123 * k) & 0xFFFFFFFF for k in range(len(ctxs))]
            t0 = time.perf_counter()
            rc, counts = stapi.run_batch(lib, ctxs, seeds)
            sync()
            first_s = time.perf_counter() - t0
            if rc != 1:
                raise RuntimeError("st_amd_sim_run_batch failed: " +
                                   "; ".join(m for c in ctxs for m in stapi.messages(lib, c)[-1:]))
            t0 = time.perf_counter()
            for _ in range(reps):
                stapi.run_batch(lib, ctxs, seeds)
            sync()
            batch_s = (time.perf_counter() - t0) / reps
            tm = (C.c_double * 70)()
            lib.dll.st_amd_batch_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
            lib.dll.st_amd_batch_timing(ctxs[0], tm, 70)
            tm = list(tm)
            t0 = time.perf_counter()
            for c, sd in zip(ctxs, seeds):
                _run(lib, c, sd)
            sync()
            sep_s = time.perf_counter() - t0
            tune = (C.c_int * 48)()
            ntuned = lib.dll.st_amd_batch_tuning(tune, 48)
        finally:
            for c in ctxs:
                lib.st_free_context(c)
    hits = sum(sc.raycount for sc in scenes)
    n_field = sum(1 for sc in scenes if sc.stages[0].elements[0].surface == "s")
    return {
        "workload": f"{n_scenes} scenes (seed 2026): {n_field} tower fields of 200-5,000 spherical heliostats, "
                    f"{n_scenes - n_field} dishes of order-6 Zernike facets and torus sections; suns over a year "
                    f"at 35 N; 1e5-2e6 ray hits each; one st_amd_sim_run_batch call per step",
        "precision": f"fp{bits}", "ray_hits": hits, "value": hits / batch_s, "unit": "ray hits/s",
        "ms_per_step": batch_s * 1e3, "first_batch_ms": first_s * 1e3, "setup_s": setup_s,
        "separate_runs_ms": sep_s * 1e3, "batch_over_separate": sep_s / batch_s,
        "records": int(sum(counts)), "sun_rays": int(tm[3]),
        "kernel_ms": {"sun_mask": tm[0], "sunshape": tm[1], "trace_stage0": tm[6], "trace_stage1": tm[7]},
        "launches": {"sun_mask": int(tm[4]), "trace_stage0": int(tm[22]), "trace_stage1": int(tm[23])},
        "autotuned": [{"class": tune[3 * i], "tile": tune[3 * i + 1], "waves": tune[3 * i + 2]}
                      for i in range(min(ntuned, 16))],
    }


def mlp_micro(batch=512, steps=200, rows=20_000):
    """The surrogate network's training step alone: `steps` graph-replayed
    Adam steps of the 8 -> 6 x 256 -> 4096 network on random data, timed
    with HIP events on torch's stream (where they launch)."""
    import torch
    from soltrace_amd.surrogate import Mlp
    net = Mlp(_lib(), 8, 256, 6, 4096, max_batch=1024, seed=1)
    X = torch.rand(rows, 8, device="cuda") * 2 - 1
    T = torch.rand(rows, 4096, device="cuda").to(torch.bfloat16)
    idx = torch.randint(0, rows, (steps, batch), device="cuda", dtype=torch.int32)
    lr = torch.full((steps,), 1e-4, device="cuda")
    net.fit(X, T, idx, lr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    net.fit(X, T, idx, lr)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    macs = 8 * 256 + 5 * 256 * 256 + 256 * 4096
    tf = 3 * 2 * batch * macs / (ms * 1e-3) / 1e12
    net.close()
    return {"batch": batch, "steps": steps, "train_step_ms": ms, "tflops": tf, "frac_bf16": tf / 2500.0}


def copy_calibration(mib):
    """A plain device-to-device copy of known size: what FETCH_SIZE and
    WRITE_SIZE read for it gives the factor between those counters and
    bytes moved."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=123)
    ap.add_argument("--hits", type=int, default=N_HITS)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-workloads", action="store_true", help="only the tower workload")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the flux-surrogate workload")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scene workload")
    ap.add_argument("--only", choices=["batched", "surrogate", "mlp"], default=None,
                    help="run one sub-workload alone and print its JSON (for profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 4 times (scale for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the trace path would shard by ray ranges, but no multi-GPU run is built)")
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    sync = torch.cuda.synchronize
    if args.only:
        log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
        if args.only == "batched":
            out = batch_workload(_lib(), args.seed, sync, log)
        elif args.only == "surrogate":
            out = surrogate_workload(_lib(), args.seed, log)
        else:
            out = mlp_micro()
        print(json.dumps({args.only: out}))
        return

    lib = _lib()
    scene = workloads.tower(n_helio=1000, raycount=args.hits)
    c32 = _context(lib, scene, 32)
    c64 = _context(lib, scene, 64)

    t32, tm32 = _timed_steps(lib, c32, args.steps, args.warmup, args.seed, sync)
    f32 = _flux(lib, c32)
    t64, tm64 = _timed_steps(lib, c64, args.steps, max(1, args.warmup), args.seed, sync)
    f64 = _flux(lib, c64)
    nrmse = float(np.sqrt(np.mean((f32 - f64) ** 2)) / max(1.0, f64.max()))

    # event counts of one FP32 run (a separate, instrumented run)
    lib.dll.st_amd_count_events(c32, 1)
    _run(lib, c32, args.seed)
    ev = _arr(lib, "st_amd_events", c32, 104)
    lib.dll.st_amd_count_events(c32, 0)
    # kernel times of the timed FP32 runs (mean per run)
    kms = {"sun_mask": tm32[0], "sunshape": tm32[52]}
    for s in range(len(scene.stages)):
        kms[f"trace_stage{s}"] = tm32[1 + s]
    dom = max(kms, key=kms.get)
    cw = counted_work(lambda n: workloads.tower(n_helio=1000, raycount=n), 4000, args.seed)
    roof = None
    if cw is not None:
        work = kernel_work(cw, ev, args.hits, tm32)
        w = work[dom]
        launches = max(1.0, w["launches"])
        avg_ms = kms[dom] / launches
        flops_per_launch = w["flops"] / launches
        achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
        units_per_launch = w["units"] / launches
        traffic = PMC_BYTES_PER_UNIT[dom] * units_per_launch if dom in PMC_BYTES_PER_UNIT else None
        roof = {
            "bound": "valu",
            "pipe": "vector ALU, FP32 and FP64 (no contraction, so no MFMA; HBM traffic is a few percent of "
                    "8 TB/s). peak is the published FP32 vector rate, which needs packed FP32; the kernels "
                    "issue non-packed FP32 and FP64, whose rate is the 78.6 TF FP64 peak (frac_nonpacked)",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "frac_nonpacked": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "traffic_source": PMC_SOURCE,
            "flops_per_launch": flops_per_launch,
            "avg_launch_ms": avg_ms,
            "flops_per_unit": {k: round(v["flops"], 1) for k, v in cw.items() if isinstance(v, dict)},
            "transcendentals_per_unit": {k: round(v["trans"], 2) for k, v in cw.items() if isinstance(v, dict)},
            "work_sample": cw["sample"],
            "kernels": {k: {"ms": round(kms.get(k, 0.0), 3), "launches": v["launches"], "units": v["units"],
                            "tflops": (v["flops"] / (kms[k] * 1e-3) / 1e12) if kms.get(k) else None}
                        for k, v in work.items()},
        }

    hits_s = args.hits / t32
    line = {
        "metric": METRIC,
        "value": hits_s,
        "unit": "ray hits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "tower: 1,000 spherical 10x10 m heliostats (radially staggered, r<=1 km), flat 15x15 m "
                        "receiver at 150 m, Buie CSR 0.1 sun, 1.5 mrad slope error, sun el 45 / az 180, "
                        f"{args.hits:.0e} ray hits, 100x100 flux map on the GPU",
            "ray_hits": args.hits,
            "sun_rays_per_step": int(tm32[19]),
            "records_per_step": int(sum(tm32[36:52])),
            "precision": "FP32 bulk ray algebra, FP64 frame transforms",
        },
        "fp64": {"value": args.hits / t64, "ms_per_step": t64 * 1e3},
        "fp32_over_fp64": t64 / t32,
        "flux_nrmse_fp32_vs_fp64": nrmse,
        "kernel_ms": {k: round(v, 3) for k, v in kms.items()},
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_workloads:
        line["end_to_end"] = _end_to_end(lib, c32, args.seed, sync)
    lib.st_free_context(c32)
    lib.st_free_context(c64)
    if not args.no_workloads:
        line["workloads"] = {
            "trough_fp64": dict(_workload_rate(lib, workloads.trough(raycount=1_000_000), 64, 5, args.seed, sync),
                                workload="parabolic trough 5.77 x 12 m, f 1.71 m, 70 mm tube, pillbox 4.65 mrad, "
                                         "1e6 ray hits"),
            "trough_fp32": _workload_rate(lib, workloads.trough(raycount=1_000_000), 32, 5, args.seed, sync),
            "faceted_fp32": dict(_workload_rate(lib, workloads.faceted_field(n_helio=10_000, raycount=50_000_000),
                                                32, 3, args.seed, sync),
                                 workload="10,000 heliostats x 4 flat 5 x 5 m facets, cylindrical receiver "
                                          "R 8.5 m x 20 m, Gaussian sun 2.5 mrad, 5e7 ray hits"),
            "faceted_fp64": _workload_rate(lib, workloads.faceted_field(n_helio=10_000, raycount=50_000_000), 64,
                                           2, args.seed, sync),
        }
    if not args.no_workloads and not args.no_batch:
        line["batched"] = batch_workload(lib, args.seed, sync, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_workloads and not args.no_surrogate:
        line["surrogate"] = surrogate_workload(lib, args.seed, lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lambda n: workloads.tower(n_helio=1000, raycount=n), args.seed)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
