"""Write tests/golden/flux/*.npz: receiver flux maps of the reference's own
FP64 trace for two workloads, the integer hit counts the GUI's flux map bins
(app/src/project.cpp:1139-1436: each ray's last record on the receiver,
absorbed only, in the receiver's frame, the tube unrolled to arc length):

* trough: workloads.trough(raycount=100000), seed 123, 50 x 50 bins on the
  absorber tube over its whole circumference and length;
* tower100: workloads.tower(n_helio=100, raycount=100000, r_max=500), seed
  1234, 100 x 100 bins over the 15 x 15 m receiver.

The records are the reference build's (oracle/_ref/coretrace_ref.so, from
`make -C oracle ref`): the script traces each scene with it and with the
oracle in the reference's stream mode, requires the two record streams to be
equal bit for bit, and bins them with the oracle's restatement of the GUI
binning (st_oracle_flux, pinned by tests/test_flux_stats.py).

Usage: python tests/golden/make_flux_golden.py"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from soltrace_amd import stapi, workloads  # noqa: E402
from harness import run_scene  # noqa: E402

REF_SO = os.path.join(ROOT, "oracle", "_ref", "coretrace_ref.so")
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "st_oracle.so")

# name -> (scene, seed, nx, ny, x range, y range)
CASES = {
    "trough": (lambda: workloads.trough(raycount=100_000), 123, 50, 50, (-6.0, 6.0), (-6.0, 6.0)),
    "tower100": (lambda: workloads.tower(n_helio=100, raycount=100_000, r_max=500.0), 1234, 100, 100,
                 (-7.5, 7.5), (-7.5, 7.5)),
}


def oracle_map(lib, ctx, nx, ny, xr, yr):
    f = lib.dll.st_oracle_flux
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 6
    f.restype = C.c_int
    vals = [C.c_double(v) for v in (*xr, *yr)]
    out = (C.c_double * (nx * ny))()
    st = (C.c_double * 3)()
    assert f(ctx, 1, 0, nx, ny, 1, 0, *[C.byref(v) for v in vals], out, st) == 1
    return np.frombuffer(out, dtype=np.float64).reshape(nx, ny).astype(np.int64)


def main():
    if not (os.path.exists(REF_SO) and os.path.exists(ORACLE_SO)):
        sys.exit("build the oracle and the reference first: make -C oracle && make -C oracle ref")
    ref, orc = stapi.StLib(REF_SO), stapi.StLib(ORACLE_SO)
    out_dir = os.path.join(HERE, "flux")
    os.makedirs(out_dir, exist_ok=True)
    for name, (make, seed, nx, ny, xr, yr) in CASES.items():
        sc = make()
        n_r, r_r, m_r, _ = run_scene(ref, sc, seed)
        n_o, r_o, m_o, grid = run_scene(orc, sc, seed, after=lambda lib, ctx: oracle_map(lib, ctx, nx, ny, xr, yr))
        assert n_r == n_o > 0, (n_r, n_o, m_r, m_o)
        for f in ("pos", "cos"):
            assert np.array_equal(getattr(r_r, f).view(np.uint64), getattr(r_o, f).view(np.uint64)), f
        for f in ("element", "stage", "raynum"):
            assert np.array_equal(getattr(r_r, f), getattr(r_o, f)), f
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), counts=grid.astype(np.int32),
                            hits=np.int64(sc.raycount), seed=np.int64(seed), xr=np.array(xr), yr=np.array(yr),
                            sun_rays=np.int64(r_r.nsunrays))
        print(name, "records", n_r, "binned", int(grid.sum()), "sun rays", r_r.nsunrays)


if __name__ == "__main__":
    main()
