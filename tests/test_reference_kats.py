"""Known answers the reference's own tests hold (CPU only).

app/deploy/api/examples/parabolic-trough-tracking-timeseries/tests/
test_intercept_factor.py:52-57 expects intercept factors 1, 0.7066204592546165
and 0 for tracking errors of 0, 1.733333e-2 rad and 2.5 deg, with the sun
overhead (nominal trough angle 0) and to the east (90).  The run is
run_pysoltrace_iterate.py in validation mode with ideal optics and
PySolTrace.run(10, False, 4): four processes, seeds 10 + 123 i passed as
c_uint16, 1e6/4 ray hits and 1e8/4 maximum rays each, each process forcing
as_power_tower (pysoltrace.py:29-31, 1599-1628); the frames are
concatenated and the intercept factor is calc_intercept_factor of
st_processing_functions.py:146-159.

1 and 0 follow from the geometry and are checked exactly.  The middle
value is 696720/985989 reduced, a count of that run's random stream; the
reference built here (libstdc++), and the oracle with it, count
696187/985916 = 0.706132 on the same scene and seeds, 1.1 standard
deviations of the binomial noise away.  Neither libc++'s nor MSVC's
default engine (minstd_rand, mt19937) reproduces it either, so the test
holds the oracle to it within 3 sigma, and to the reference build's exact
counts where that build is present."""
import math

import numpy as np
import pytest

from soltrace_amd.scene import Element, Optic, OpticFace, Scene, Stage
from harness import run_scene

F, D_ABS, L, W = 1.49, 0.07, 12.0, 5.0


def _trough(trough_angle_deg, sun_xyz, hits):
    # run_pysoltrace_iterate.py:191-275 with pysoltrace's defaults
    # (element aim (0,0,0) before aim.z is set, sun 'p' 4.65 mrad)
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=sun_xyz, raycount=hits, maxcount=int(1e8) // 4,
              dyn_group=True, sunshape_err=False, optics_err=False)
    s.optics = [Optic("Reflector", OpticFace(dist="g", refl=1.0), OpticFace(dist="g", refl=1.0)),
                Optic("Absorber", OpticFace(dist="g", refl=0.0), OpticFace(dist="g", refl=0.0))]
    trough = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="r", ap_params=(W, L), surface="p",
                     surf_params=(1.0 / (2.0 * F), 1.0 / (2.0 * math.inf)), optic="Reflector")
    tube = Element(xyz=(0, 0, F - D_ABS / 2.0), aim=(0, 0, 2.0 * F), aperture="l", ap_params=(0.0, 0.0, L),
                   surface="t", surf_params=(1.0 / (D_ABS / 2.0),), optic="Absorber")
    a = math.radians(trough_angle_deg)
    s.stages = [Stage(xyz=(0, 0, 0), aim=(math.sin(a), 0.0, math.cos(a)), elements=[trough, tube])]
    return s


def _intercept_factor(lib, trough_angle_deg, sun_xyz, total_hits=1_000_000, nproc=4, seed=10):
    coll = rcvr = 0
    for i in range(nproc):
        hits = total_hits // nproc + (total_hits % nproc if i == 0 else 0)
        n, r, msgs, _ = run_scene(lib, _trough(trough_angle_deg, sun_xyz, hits), (seed + 123 * i) & 0xFFFF)
        assert n > 0, msgs
        st1 = r.stage == 1
        coll += int((st1 & (r.element == 1)).sum())
        neg = r.element[r.element < 0]
        if neg.size:
            rcvr += int((st1 & (r.element == neg.min()) & (r.cos[:, 2] > 0.0)).sum())
    return rcvr / coll


ERR_RAD_DEG = float(np.degrees(1.733333e-02))
KAT = 0.7066204592546165


@pytest.mark.parametrize("nominal,sun_xyz", [(0.0, (0.0, 0.0, 100.0)), (90.0, (100.0, 0.0, 0.0))])
@pytest.mark.parametrize("error,want", [(0.0, 1.0), (2.5, 0.0)])
def test_intercept_factor_geometric_kats(oracle_lib, nominal, sun_xyz, error, want):
    assert _intercept_factor(oracle_lib, nominal + error, sun_xyz) == want


def _counts(lib, trough_angle_deg, sun_xyz):
    out = []
    for i in range(4):
        n, r, msgs, _ = run_scene(lib, _trough(trough_angle_deg, sun_xyz, 250000), (10 + 123 * i) & 0xFFFF)
        assert n > 0, msgs
        out.append((n, r.nsunrays, int((r.element == 1).sum()), int((r.element == -2).sum())))
    return out


@pytest.mark.parametrize("nominal,sun_xyz", [(0.0, (0.0, 0.0, 100.0)), (90.0, (100.0, 0.0, 0.0))])
def test_intercept_factor_statistical_kat(oracle_lib, nominal, sun_xyz):
    got = _intercept_factor(oracle_lib, nominal + ERR_RAD_DEG, sun_xyz)
    sigma = math.sqrt(KAT * (1 - KAT) / 985989)
    assert abs(got - KAT) < 3 * sigma, (got, KAT, sigma)


def test_intercept_factor_counts_equal_reference_build(oracle_lib, ref_lib):
    assert _counts(oracle_lib, ERR_RAD_DEG, (0.0, 0.0, 100.0)) == _counts(ref_lib, ERR_RAD_DEG, (0.0, 0.0, 100.0))
