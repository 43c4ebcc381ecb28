"""Build the HIP backend in-tree for gfx950.

Produces ``soltrace_amd/coretrace_api.so`` (the name the reference's callers
load, coretrace/CMakeLists.txt:89-94, no ``lib`` prefix) and the static
``soltrace_amd/libcoretrace.a`` (coretrace/CMakeLists.txt:105-110).

Usage: ``python -m soltrace_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("SOLTRACE_AMD_ARCH", "gfx950")

SOURCES = ["st_trace_f64.hip", "st_trace_f32.hip", "st_compact.hip", "st_stats.hip", "st_scene.cpp", "st_context.cpp",
           "st_batch.cpp", "st_mlp.hip"]
# The FP64 trace keeps separate multiply/add roundings (no FMA contraction),
# the arithmetic the reference's CPU build performs.
EXTRA = {"st_trace_f64.hip": ["-ffp-contract=off"], "st_trace_f32.hip": ["-ffp-contract=fast"],
         "st_stats.hip": ["-ffp-contract=off"]}

# -packed-fp32-ops off: packed FP32 forms are excluded from every kernel.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Wno-unused-command-line-argument",
    "-I" + CSRC,
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP backend)")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


# Host C++ built into the library beside the C API: the .stinput reader and
# the ray-file writers (they call only stapi.h).
STRACE = os.path.join(HERE, "strace")
HOST_SOURCES = ["stinput.cpp", "stinput_capi.cpp"]
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build(verbose: bool = False, only=None) -> str:
    """Compile every source (in parallel) and link.  `only` (a list of
    source names) recompiles just those and relinks with the other objects
    of an earlier build."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append((src, [cc, *lang, *HIP_FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]))
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmds.append((src, ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-c",
                           os.path.join(STRACE, src), "-o", obj]))
        objs.append(obj)
    todo = [c for s, c in cmds if only is None or s in only or not os.path.exists(c[-1])]
    if verbose:
        for c in todo:
            print(" ".join(c), flush=True)
    jobs = max(1, min(len(todo), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        for _ in ex.map(_run, todo):
            pass
    so = os.path.join(HERE, "coretrace_api.so")
    tmp = so + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, so)
    lib = os.path.join(HERE, "libcoretrace.a")
    if os.path.exists(lib):
        os.remove(lib)
    _run(["ar", "rcs", lib, *objs])
    build_strace(so, verbose)
    return so


def build_strace(so: str, verbose: bool = False) -> str:
    """The strace command line (soltrace_amd/strace), host C++ over the
    C API; its reader and writer are those built into the shared library."""
    src = os.path.join(HERE, "strace")
    exe = os.path.join(src, "strace")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, os.path.join(src, "main.cpp"),
           "-L" + os.path.dirname(so), "-l:" + os.path.basename(so),
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-o", exe + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
