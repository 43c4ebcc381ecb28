"""Checks fixed by the physics and geometry alone, usable on any stapi
library (the oracle on the CPU, the HIP library on a GPU).  Each check
builds a small scene, runs it and asserts what the mathematics fixes:

* reflection law on a tilted flat mirror, exact hit plane;
* a parabolic dish sends collimated light through its focus;
* sphere hits lie on the sphere and reflect about the radial normal;
* refraction obeys Snell's law and splits by the average Fresnel
  reflectance;
* the ideal trough intercept factor is 1 and 0 at 2.5 deg tracking error
  (the geometry of app/deploy/api/examples/parabolic-trough-tracking-
  timeseries/run_pysoltrace_iterate.py, validation mode);
* roulette: the absorbed fraction equals 1 - reflectivity;
* sunshape moments: pillbox E[theta^2] = R^2/2 within the disk, Gaussian
  E[theta^2] = 2 sigma^2, the Buie table's polar-angle CDF;
* direction cosines are unit vectors.

`tol` scales the geometric tolerances (1 for FP64 arithmetic)."""
import math

import numpy as np

from soltrace_amd.scene import Element, Optic, OpticFace, Scene, Stage
from soltrace_amd import workloads
from harness import run_scene


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def _run(lib, scene, seed, setup):
    n, r, msgs, _ = run_scene(lib, scene, seed, setup=setup)
    assert n > 0, msgs
    return r


def _mirror_optics(refl=1.0):
    return [Optic("m", OpticFace(dist="g", refl=refl), OpticFace(dist="g", refl=refl)),
            Optic("a", OpticFace(dist="g", refl=0.0), OpticFace(dist="g", refl=0.0))]


def reflection_law(lib, setup=None, tol=1.0):
    sun = np.array([30.0, -20.0, 100.0])
    d = -unit(sun)
    n = unit([0.2, 0.1, 1.0])
    r = d - 2.0 * d.dot(n) * n
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=tuple(sun), raycount=3000, sunshape_err=False,
              optics_err=False)
    s.optics = _mirror_optics()
    mirror = Element(xyz=(0, 0, 0), aim=tuple(n), aperture="r", ap_params=(4.0, 4.0), surface="f", optic="m")
    tgt = 20.0 * r
    absorber = Element(xyz=tuple(tgt), aim=(0, 0, 0), aperture="r", ap_params=(40.0, 40.0), surface="f",
                       optic="a")
    s.stages = [Stage(elements=[mirror]), Stage(elements=[absorber])]
    res = _run(lib, s, 7, setup)
    s1, s2 = res.stage == 1, res.stage == 2
    assert (res.element[s1] == 1).all() and s2.sum() == s1.sum()
    assert (res.element[s2] == -1).all()
    np.testing.assert_allclose(res.cos[s1], np.broadcast_to(d, res.cos[s1].shape), atol=1e-12 * tol)
    np.testing.assert_allclose(res.pos[s1] @ n, 0.0, atol=1e-9 * tol)
    np.testing.assert_allclose(res.cos[s2], np.broadcast_to(r, res.cos[s2].shape), atol=1e-12 * tol)


def parabolic_focus(lib, setup=None, tol=1.0):
    f = 2.0
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=(0.0, 0.0, 100.0), raycount=3000, sunshape_err=False,
              optics_err=False)
    s.optics = _mirror_optics()
    dish = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="c", ap_params=(3.0,), surface="p",
                   surf_params=(1 / (2 * f), 1 / (2 * f)), optic="m")
    rec = Element(xyz=(0, 0, f), aim=(0, 0, 0), aperture="c", ap_params=(0.5,), surface="f", optic="a")
    s.stages = [Stage(elements=[dish]), Stage(elements=[rec])]
    res = _run(lib, s, 3, setup)
    p1 = res.pos[res.stage == 1]
    np.testing.assert_allclose(p1[:, 2], (p1[:, 0] ** 2 + p1[:, 1] ** 2) / (4 * f), atol=1e-9 * tol)
    p2 = res.pos[res.stage == 2]
    assert len(p2) == len(p1)
    assert np.hypot(p2[:, 0], p2[:, 1]).max() < 1e-9 * tol
    np.testing.assert_allclose(p2[:, 2], f, atol=1e-9 * tol)


def sphere_reflection(lib, setup=None, tol=1.0):
    R = 5.0
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=(10.0, 5.0, 100.0), raycount=3000, sunshape_err=False,
              optics_err=False)
    s.optics = _mirror_optics()
    dish = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="c", ap_params=(4.0,), surface="s",
                   surf_params=(1 / R,), optic="m")
    rec = Element(xyz=(0, 0, 1.0), aim=(0, 0, 0), aperture="r", ap_params=(20.0, 20.0), surface="f",
                  optic="a")
    s.stages = [Stage(elements=[dish]), Stage(elements=[rec])]
    res = _run(lib, s, 5, setup)
    m1, m2 = res.stage == 1, res.stage == 2
    p1 = res.pos[m1]
    np.testing.assert_allclose(np.linalg.norm(p1 - [0, 0, R], axis=1), R, atol=1e-9 * tol)
    # reflected direction about the radial normal, matched by ray number
    d = res.cos[m1]
    nrm = (np.array([0, 0, R]) - p1) / R
    refl = d - 2.0 * np.sum(d * nrm, axis=1)[:, None] * nrm
    order2 = {int(k): i for i, k in enumerate(res.raynum[m2])}
    idx = [order2[int(k)] for k in res.raynum[m1]]
    np.testing.assert_allclose(res.cos[m2][idx], refl, atol=1e-9 * tol)


def refraction_snell_fresnel(lib, setup=None, tol=1.0):
    n1, n2, inc = 1.0, 1.5, math.radians(30.0)
    sun = 100.0 * np.array([math.sin(inc), 0.0, math.cos(inc)])
    d = -unit(sun)
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=tuple(sun), raycount=20000, sunshape_err=False,
              optics_err=False)
    s.optics = [Optic("glass", OpticFace(dist="g", rreal=n1, trans=1.0), OpticFace(dist="g", rreal=n2, trans=1.0)),
                Optic("a", OpticFace(dist="g", refl=0.0), OpticFace(dist="g", refl=0.0))]
    face = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="r", ap_params=(4.0, 4.0), surface="f", optic="glass",
                   interaction=1)
    below = Element(xyz=(0, 0, -5), aim=(0, 0, 0), aperture="r", ap_params=(100.0, 100.0), surface="f", optic="a")
    above = Element(xyz=(0, 0, 5), aim=(0, 0, 0), aperture="r", ap_params=(100.0, 100.0), surface="f", optic="a")
    s.stages = [Stage(elements=[face]), Stage(elements=[below, above])]
    res = _run(lib, s, 11, setup)
    m2 = res.stage == 2
    down = res.cos[m2][np.abs(res.element[m2]) == 1]
    up = res.cos[m2][np.abs(res.element[m2]) == 2]
    assert len(down) + len(up) == (res.stage == 1).sum()
    # Snell: tangential component scaled by n1/n2, still a unit vector
    sin_t = math.sin(inc) * n1 / n2
    want_t = np.array([-sin_t, 0.0, -math.sqrt(1.0 - sin_t ** 2)])
    np.testing.assert_allclose(down, np.broadcast_to(want_t, down.shape), atol=2e-5)
    want_r = d * np.array([1.0, 1.0, -1.0])
    np.testing.assert_allclose(up, np.broadcast_to(want_r, up.shape), atol=1e-12 * tol)
    # average of the s and p Fresnel reflectances (textbook form)
    ci, ct = math.cos(inc), math.sqrt(1.0 - sin_t ** 2)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    R = 0.5 * (rs ** 2 + rp ** 2)
    n = len(down) + len(up)
    assert abs(len(up) / n - R) < 4.0 * math.sqrt(R * (1 - R) / n)


def _ideal_trough(tilt_deg, n=20000):
    # run_pysoltrace_iterate.py validation geometry, single stage, ideal optics
    f, d, L, W = 1.49, 0.07, 12.0, 5.0
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=(0.0, 0.0, 100.0), raycount=n, sunshape_err=False,
              optics_err=False)
    s.optics = [Optic("Reflector", OpticFace(dist="g", refl=1.0), OpticFace(dist="g", refl=1.0)),
                Optic("Absorber", OpticFace(dist="g", refl=0.0), OpticFace(dist="g", refl=0.0))]
    trough = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="r", ap_params=(W, L), surface="p",
                     surf_params=(1 / (2 * f), 0.0), optic="Reflector")
    tube = Element(xyz=(0, 0, f - d / 2), aim=(0, 0, 2 * f), aperture="l", ap_params=(0.0, 0.0, L), surface="t",
                   surf_params=(2 / d,), optic="Absorber")
    t = math.radians(tilt_deg)
    s.stages = [Stage(aim=(math.sin(t), 0.0, math.cos(t)), elements=[trough, tube])]
    return s


def _intercept_factor(res):
    # calc_intercept_factor of the example's st_processing_functions.py
    coll = ((res.stage == 1) & (res.element == 1)).sum()
    rcvr = ((res.stage == 1) & (res.element == -2) & (res.cos[:, 2] > 0)).sum()
    return rcvr / coll


def trough_intercept_factors(lib, setup=None, tol=1.0):
    assert _intercept_factor(_run(lib, _ideal_trough(0.0), 10, setup)) == 1.0
    assert _intercept_factor(_run(lib, _ideal_trough(2.5), 10, setup)) == 0.0


def roulette_fraction(lib, setup=None, tol=1.0):
    s = Scene(sun_shape="p", sun_sigma=4.65, sun_xyz=(0.0, 0.0, 100.0), raycount=20000, sunshape_err=False,
              optics_err=False)
    s.optics = _mirror_optics(refl=0.7)
    s.stages = [Stage(elements=[Element(aperture="r", ap_params=(2.0, 2.0), surface="f", optic="m")])]
    res = _run(lib, s, 13, setup)
    frac = (res.element < 0).mean()
    assert abs(frac - 0.3) < 4.0 * math.sqrt(0.3 * 0.7 / len(res))


def _sun_spread(lib, setup, shape, sigma, table=None, n=20000):
    # flat mirror facing the sun at the zenith: the reflected ray's angle to
    # +z is the sunshape's polar angle
    s = Scene(sun_shape=shape, sun_sigma=sigma, sun_xyz=(0.0, 0.0, 100.0), raycount=n, sunshape_err=True,
              optics_err=False, sun_table=table)
    s.optics = _mirror_optics()
    mirror = Element(aperture="r", ap_params=(2.0, 2.0), surface="f", optic="m")
    top = Element(xyz=(0, 0, 10), aim=(0, 0, 0), aperture="r", ap_params=(20.0, 20.0), surface="f", optic="a")
    s.stages = [Stage(elements=[mirror]), Stage(elements=[top])]
    res = _run(lib, s, 17, setup)
    c = res.cos[res.stage == 2]
    assert len(c) == n
    return np.arccos(np.clip(c[:, 2], -1.0, 1.0))


def sunshape_moments(lib, setup=None, tol=1.0):
    R = 4.65e-3
    th = _sun_spread(lib, setup, "p", 4.65)
    assert th.max() <= R * (1 + 1e-6)
    m = np.mean(th ** 2)
    assert abs(m / (R * R / 2) - 1) < 4 * math.sqrt(1 / 3) / math.sqrt(len(th))  # var(u^2)/E^2 = 1/3 for u^2 ~ U(0,1)
    sg = 2.5e-3
    th = _sun_spread(lib, setup, "g", 2.5)
    m = np.mean(th ** 2)
    assert abs(m / (2 * sg * sg) - 1) < 4.0 / math.sqrt(len(th))  # theta^2 / 2 sigma^2 ~ Exp(1)


def buie_polar_cdf(lib, setup=None, tol=1.0):
    ang, inten = workloads.buie_profile(0.1)
    th = _sun_spread(lib, setup, "d", 0.0, table=(ang, inten)) * 1000.0  # mrad
    # the reference samples (tx, ty) over the disk of radius max(ang) with
    # acceptance prof(theta)/max: polar density proportional to prof * theta
    grid = np.linspace(0.0, ang[-1], 200001)
    prof = np.interp(grid, ang, inten)
    dens = prof * grid
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    th = np.sort(th)
    emp = np.arange(1, len(th) + 1) / len(th)
    ks = np.max(np.abs(emp - np.interp(th, grid, cdf)))
    assert ks < 1.63 / math.sqrt(len(th)), ks  # p = 0.01


def unit_cosines(lib, setup=None, tol=1.0):
    res = _run(lib, workloads.trough(raycount=5000), 21, setup)
    np.testing.assert_allclose(np.linalg.norm(res.cos, axis=1), 1.0, atol=1e-12 * tol)


ALL = [reflection_law, parabolic_focus, sphere_reflection, refraction_snell_fresnel, trough_intercept_factors,
       roulette_fraction, sunshape_moments, buie_polar_cdf, unit_cosines]
