"""The north-star flux criterion: the GPU's production flux map (per-ray
Philox streams, device binning) agrees with the reference's own FP64 trace
within Monte Carlo noise.  The reference maps are fixtures written by
tests/golden/make_flux_golden.py from the reference build's records (the
GUI's binning, app/src/project.cpp:1139-1436).  The two runs draw
independent random numbers, so they are compared as two samples:

* per-bin two-sample chi-square over bins holding any count, on the full
  grid and on a 5 x 5-coarsened grid (where counts are large enough for the
  chi-square law to hold well): p-value above 1e-3;
* intercepted count (hits binned per ray hit) within 3 standard errors of
  the difference of two binomial proportions."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import stats

from soltrace_amd import stapi, workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = {
    "trough": (lambda: workloads.trough(raycount=100_000), 50, 50, (-6.0, 6.0), (-6.0, 6.0)),
    "tower100": (lambda: workloads.tower(n_helio=100, raycount=100_000, r_max=500.0), 100, 100, (-7.5, 7.5),
                 (-7.5, 7.5)),
}


def chi2_two_sample(a, b):
    """Chi-square statistic and p-value of two unweighted histograms with
    different totals (bins empty in both are left out)."""
    a, b = a.astype(np.float64).ravel(), b.astype(np.float64).ravel()
    na, nb = a.sum(), b.sum()
    keep = (a + b) > 0
    a, b = a[keep], b[keep]
    x2 = np.sum((np.sqrt(nb / na) * a - np.sqrt(na / nb) * b) ** 2 / (a + b))
    dof = keep.sum() - 1
    return x2, dof, stats.chi2.sf(x2, dof)


def coarsen(m, k):
    nx, ny = m.shape
    return m[: nx // k * k, : ny // k * k].reshape(nx // k, k, ny // k, k).sum(axis=(1, 3))


def gpu_map(lib, sc, bits, seed, nx, ny, xr, yr):
    d = lib.dll
    ctx = lib.st_create_context()
    try:
        sc.apply(lib, ctx)
        assert d.st_amd_set_precision(ctx, bits) == 1
        assert d.st_amd_flux_request(ctx, 1, 0, nx, ny, 1, xr[0], xr[1], yr[0], yr[1]) == 1
        assert stapi.run(lib, ctx, seed) > 0, stapi.messages(lib, ctx)
        out = (C.c_uint * (nx * ny))()
        assert d.st_amd_flux_result(ctx, out, None) == nx * ny
        return np.frombuffer(out, dtype=np.uint32).reshape(nx, ny).astype(np.int64)
    finally:
        lib.st_free_context(ctx)


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_flux_map_agrees_with_the_reference_trace(hip_lib, name, bits):
    make, nx, ny, xr, yr = SCENES[name]
    g = np.load(os.path.join(HERE, "golden", "flux", name + ".npz"))
    ref, hits = g["counts"].astype(np.int64), int(g["hits"])
    sc = make()
    assert sc.raycount == hits
    gpu = gpu_map(hip_lib, sc, bits, 777, nx, ny, xr, yr)
    for k in (1, 5):
        x2, dof, p = chi2_two_sample(coarsen(gpu, k), coarsen(ref, k))
        print(f"{name} FP{bits} {k}x{k}: chi2 {x2:.1f} / {dof} dof, p = {p:.3g}")
        assert p > 1e-3, (k, x2, dof, p)
    pa, pb = gpu.sum() / hits, ref.sum() / hits
    se = np.sqrt(pa * (1 - pa) / hits + pb * (1 - pb) / hits)
    print(f"intercepted {gpu.sum()} vs {ref.sum()} ({(pa - pb) / se:+.2f} sigma)")
    assert abs(pa - pb) <= 3.0 * se, (gpu.sum(), ref.sum())
