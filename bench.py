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
FP64 speed ratio and the flux-map NRMSE = RMS(F32 - F64) / max(F64).

`cpu_baseline` times the reference's own coretrace, compiled from source in
this repository's oracle/ recipe (oracle/_ref/coretrace_ref.so), on the
host's cores with the GUI's multithreaded scheme (one context per thread,
rays split evenly, seeds seed, seed+123, seed+369, ...), on a bounded sample
of the same scene.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector peak
N_HITS = 10_000_000
NX = NY = 100

# Hand-counted FP work per kernel event (FMA = 2 flops, add/mul/div/sqrt = 1,
# transcendentals 1 each), from the device code in soltrace_amd/csrc.
FLOPS = {
    "sun_ray": 110,        # sun-plane point, sun->stage->global->stage transforms, grid cell
    "stage_ray": 33,       # global -> stage transform of an entering ray
    "test_sphere": 115,    # element transform + nudge + rebase + float quadric roots + aperture
    "test_flat": 50,       # element transform + nudge + start-plane point + aperture
    "bvh_node": 12,        # float slab test
    "uniform": 5,          # conversion + rejection-loop arithmetic per uniform drawn in the kernel
    "record": 420,         # double hit finish, record transform, efficiency, 2-3 perturbations, reflection
    "sampler_try": 14,     # sunshape screen: proposal, radius, bin, envelope test (Philox is integer work)
}

# HBM bytes per unit (ray) of the dominant kernels on this workload, from
# rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script, scaled
# by what the counters read for a plain 1 GiB device copy (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0); profiles/hbm_traffic_tower1000.md has the runs.
PMC_BYTES_PER_UNIT = {
    "trace_stage0": 76.4 + 277.8,
    "sunshape": 8.6 + 16.7,
    "sun_mask": 0.14 + 0.13,
}


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
    for _ in range(warmup):
        _run(lib, ctx, seed)
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        _run(lib, ctx, seed)
    sync()
    return (time.perf_counter() - t0) / steps


def _flop_model(scene, ev, sampler_units):
    """Algorithmic flops of each kernel of one run from its event counts."""
    launches_mask = ev[0]
    launches_trace = ev[1:17]
    base = 17
    mask = ev[base:base + 5]
    per_stage = [ev[base + 5 * (1 + s): base + 5 * (2 + s)] for s in range(16)]
    kinds = [s.elements[0].surface if s.elements else "f" for s in scene.stages]
    def test_cost(k):
        return FLOPS["test_sphere"] if k in "spot" else FLOPS["test_flat"]
    out = {}
    out["sun_mask"] = dict(launches=launches_mask,
                           flops=mask[0] * FLOPS["sun_ray"] + mask[1] * test_cost(kinds[0]) +
                           mask[3] * FLOPS["uniform"], units=mask[0])
    for s in range(len(scene.stages)):
        r, tests, nodes, unif, recs = per_stage[s]
        fl = (r * (FLOPS["sun_ray"] if s == 0 else FLOPS["stage_ray"]) + tests * test_cost(kinds[s]) +
              nodes * FLOPS["bvh_node"] + unif * FLOPS["uniform"] + recs * FLOPS["record"])
        out[f"trace_stage{s}"] = dict(launches=launches_trace[s], flops=fl, units=r, tests=tests,
                                      nodes=nodes, uniforms=unif, records=recs)
    if len(ev) >= 104 and ev[102] > 0:
        out["sunshape"] = dict(launches=launches_trace[0], flops=ev[102] * FLOPS["sampler_try"],
                               units=sampler_units, tries_evaluated=ev[102], tries_needed=ev[103])
    return out


def cpu_baseline(scene_fn, seed, target_s=12.0):
    """The reference build on the host cores, GUI threading scheme
    (app/src/trace.cpp:709-729), on a bounded sample of the same scene."""
    so = os.path.join(ROOT, "oracle", "_ref", "coretrace_ref.so")
    kind = "reference"
    if not os.path.exists(so):
        so = os.path.join(ROOT, "oracle", "_build", "st_oracle.so")
        kind = "port"
    if not os.path.exists(so):
        return None
    ref = stapi.StLib(so)
    threads = max(1, min(16, os.cpu_count() or 1))
    # calibrate on one thread
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
    ctxs, scenes = [], []
    seeds = []
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
            "sample": f"{threads} contexts x {per_thread} ray hits of the same 1,000-heliostat scene "
                      f"(dynamic grouping on), {dt:.1f} s wall; single-thread calibration {rate1:.0f} hits/s"}


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

    lib = _lib()
    scene = workloads.tower(n_helio=1000, raycount=args.hits)
    c32 = _context(lib, scene, 32)
    c64 = _context(lib, scene, 64)

    t32 = _timed_steps(lib, c32, args.steps, args.warmup, args.seed, sync)
    tm32 = _arr(lib, "st_amd_timing", c32, 53)
    f32 = _flux(lib, c32)
    t64 = _timed_steps(lib, c64, max(1, args.steps // 2), 1, args.seed, sync)
    tm64 = _arr(lib, "st_amd_timing", c64, 53)
    f64 = _flux(lib, c64)
    nrmse = float(np.sqrt(np.mean((f32 - f64) ** 2)) / max(1.0, f64.max()))

    # event counts of one FP32 run (a separate, instrumented run)
    lib.dll.st_amd_count_events(c32, 1)
    _run(lib, c32, args.seed)
    ev = _arr(lib, "st_amd_events", c32, 104)
    lib.dll.st_amd_count_events(c32, 0)
    model = _flop_model(scene, ev, sampler_units=tm32[20])
    # dominant kernel of the timed FP32 runs
    kms = {"sun_mask": tm32[0], "sunshape": tm32[52]}
    for s in range(len(scene.stages)):
        kms[f"trace_stage{s}"] = tm32[1 + s]
    dom = max((k for k in kms if k in model), key=kms.get)
    launches = max(1.0, model[dom]["launches"])
    avg_ms = kms[dom] / launches
    flops_per_launch = model[dom]["flops"] / launches
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    units_per_launch = model[dom]["units"] / launches
    traffic = PMC_BYTES_PER_UNIT[dom] * units_per_launch if dom in PMC_BYTES_PER_UNIT else None

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
        "roofline": {
            "bound": "mfma",
            "pipe": "fp32 vector ALU (the ray algebra is not a contraction; packed FP32 is not used, "
                    "so the reachable FP32 FMA rate equals the 78.6 TF FP64 rate)",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_gbps": (traffic / (avg_ms * 1e-3) / 1e9) if traffic else None,
            "flops_per_launch": flops_per_launch,
            "avg_launch_ms": avg_ms,
            "events_per_run": {k: {kk: vv for kk, vv in v.items() if kk != "flops"} for k, v in model.items()},
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lambda n: workloads.tower(n_helio=1000, raycount=n), args.seed)
    lib.st_free_context(c32)
    lib.st_free_context(c64)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
