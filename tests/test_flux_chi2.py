"""The two-sample flux test of tests/test_gpu_flux_chi2.py, run on the CPU
with the oracle: the oracle on per-ray Philox streams (the GPU's draws)
against the reference trace's fixture maps must pass, and the same oracle
with a planted mistake (the heliostats' slope error doubled, or the tube
moved 5 mm off the focal line) must fail -- so the test both accepts
independent Monte Carlo noise and rejects a physics error of that size."""
import ctypes as C
import os

import numpy as np
import pytest

from harness import counter_mode, run_scene
from test_gpu_flux_chi2 import SCENES, chi2_two_sample, coarsen

HERE = os.path.dirname(os.path.abspath(__file__))


def oracle_map(oracle_lib, sc, seed, nx, ny, xr, yr):
    def after(lib, ctx):
        vals = [C.c_double(v) for v in (*xr, *yr)]
        out = (C.c_double * (nx * ny))()
        st = (C.c_double * 3)()
        assert lib.dll.st_oracle_flux(ctx, 1, 0, nx, ny, 1, 0, *[C.byref(v) for v in vals], out, st) == 1
        return np.frombuffer(out, dtype=np.float64).reshape(nx, ny).astype(np.int64)
    n, _, msgs, grid = run_scene(oracle_lib, sc, seed, setup=counter_mode, after=after)
    assert n > 0, msgs
    return grid


def verdict(m, ref, hits):
    ps = [chi2_two_sample(coarsen(m, k), coarsen(ref, k))[2] for k in (1, 5)]
    pa, pb = m.sum() / hits, ref.sum() / hits
    se = np.sqrt(pa * (1 - pa) / hits + pb * (1 - pb) / hits)
    return min(ps) > 1e-3 and abs(pa - pb) <= 3.0 * se


@pytest.mark.parametrize("name", sorted(SCENES))
def test_counter_streams_agree_with_the_reference_trace(oracle_lib, name):
    make, nx, ny, xr, yr = SCENES[name]
    g = np.load(os.path.join(HERE, "golden", "flux", name + ".npz"))
    assert verdict(oracle_map(oracle_lib, make(), 777, nx, ny, xr, yr), g["counts"], int(g["hits"]))


def test_planted_mistakes_fail(oracle_lib):
    make, nx, ny, xr, yr = SCENES["trough"]
    g = np.load(os.path.join(HERE, "golden", "flux", "trough.npz"))
    sc = make()
    sc.optics[0].front.slope *= 2.0
    assert not verdict(oracle_map(oracle_lib, sc, 777, nx, ny, xr, yr), g["counts"], int(g["hits"]))
    sc = make()
    tube = sc.stages[1].elements[0]
    tube.xyz = (tube.xyz[0], tube.xyz[1], tube.xyz[2] + 0.005)
    assert not verdict(oracle_map(oracle_lib, sc, 777, nx, ny, xr, yr), g["counts"], int(g["hits"]))
