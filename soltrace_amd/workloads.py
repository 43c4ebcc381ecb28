"""Synthetic workloads: the trough, tower and faceted-field scenes the
benchmarks and parity tests run, built only from their stated parameters.

* :func:`trough` -- parabolic trough, 5.77 m x 12 m aperture, f = 1.71 m,
  70 mm absorber tube on the focal line, pillbox sun 4.65 mrad (layout as in
  app/deploy/api/examples/trough-demo.py:36-66).
* :func:`tower` -- N spherically focused 10 x 10 m heliostats on a radially
  staggered layout within 1 km, flat 15 x 15 m receiver at 150 m, Buie CSR
  sun, sun at 45 deg elevation and 180 deg azimuth (aim = bisector of the sun
  and receiver vectors, as app/deploy/api/examples/tower-demo.py:49-54).
* :func:`faceted_field` -- N heliostats x 4 flat 5 x 5 m facets plus a
  cylindrical external receiver (as the reference's tower sample's receiver:
  element on the wall, aimed along +y, 't' surface with an 'l' aperture).
  Each facet is turned about its normal so its edges follow the heliostat's
  rows: the four facets tile the heliostat with 5 cm gaps and never overlap.
* :func:`buie_profile` -- the Buie et al. (2003) sunshape, tabulated for the
  'd' user sunshape (angles in mrad).
"""
from __future__ import annotations

import math
import os

import numpy as np

from .scene import Element, Optic, OpticFace, Scene, Stage

# ------------------------------------------------------------------ sun ----


def buie_profile(csr: float, n_disk: int = 48, n_aureole: int = 48, theta_max: float = 43.6):
    """Buie, Monger & Dey, Solar Energy 74 (2003) 113-122, eqs. (3)-(6):
    phi(t) = cos(0.326 t)/cos(0.308 t) for t <= 4.65 mrad,
    phi(t) = exp(kappa) t^gamma beyond, with
    kappa = 0.9 ln(13.5 chi) chi^-0.3, gamma = 2.2 ln(0.52 chi) chi^0.43 - 0.1.
    Returns (angles_mrad, intensity normalised to 1 at t = 0)."""
    chi = float(csr)
    kappa = 0.9 * math.log(13.5 * chi) * chi ** -0.3
    gamma = 2.2 * math.log(0.52 * chi) * chi ** 0.43 - 0.1
    disk = np.linspace(0.0, 4.65, n_disk)
    aur = np.geomspace(4.65 + 1e-3, theta_max, n_aureole)
    a = np.concatenate([disk, aur])
    phi = np.where(a <= 4.65, np.cos(0.326 * a) / np.cos(0.308 * a),
                   np.exp(kappa) * np.power(np.maximum(a, 1e-9), gamma))
    phi = phi / phi[0]
    return a, phi


# ---------------------------------------------------------------- trough ---


def trough(raycount: int = 1_000_000, aperture: float = 5.77, length: float = 12.0, focal: float = 1.71,
           tube_d: float = 0.070, refl: float = 0.94, slope: float = 2.0, spec: float = 0.5,
           absorptivity: float = 0.95, sun_half_angle: float = 4.65, sun_shape: str = "p") -> Scene:
    """Parabolic-cylinder mirror + absorber tube on the focal line."""
    r = tube_d / 2.0
    s = Scene(sun_shape=sun_shape, sun_sigma=sun_half_angle, sun_xyz=(0.0, 0.0, 100.0),
              raycount=raycount, maxcount=max(100 * raycount, 1000), dyn_group=False)
    s.optics = [Optic("mirror", OpticFace(dist="g", refl=refl, slope=slope, spec=spec),
                      OpticFace(dist="g", refl=0.0, slope=slope, spec=spec)),
                Optic("absorber", OpticFace(dist="g", refl=1.0 - absorptivity),
                      OpticFace(dist="g", refl=1.0 - absorptivity))]
    mirror = Element(xyz=(0, 0, 0), aim=(0, 0, 1), aperture="l",
                     ap_params=(-aperture / 2, aperture / 2, length), surface="p",
                     surf_params=(1.0 / (2.0 * focal), 0.0), optic="mirror")
    tube = Element(xyz=(0, 0, focal - r), aim=(0, 0, 100), aperture="l", ap_params=(0, 0, length),
                   surface="t", surf_params=(1.0 / r,), optic="absorber")
    s.stages = [Stage(elements=[mirror]), Stage(elements=[tube])]
    return s


# ----------------------------------------------------------------- tower ---


def staggered_layout(n: int, r_min: float, r_max: float, pitch: float,
                     north_only: bool = False) -> np.ndarray:
    """Radially staggered rings: heliostats spaced ``pitch`` along each ring,
    odd rings rotated by half a pitch, ring spacing chosen so that ``n``
    positions fill [r_min, r_max] (the last ring may be partial)."""
    span = math.pi if north_only else 2 * math.pi
    dr = span * (r_max ** 2 - r_min ** 2) / (2.0 * pitch * n)
    pts = []
    k = 0
    while len(pts) < n:
        r = r_min + (k + 0.5) * dr
        m = max(1, int(span * r / pitch))
        off = 0.5 if (k % 2) else 0.0
        for j in range(m):
            a = (j + off) / m * span
            pts.append((r * math.cos(a), r * math.sin(a)))
        k += 1
    return np.array(pts[:n], dtype=np.float64)


def sun_vector(elev_deg: float, azim_deg: float) -> np.ndarray:
    """Unit vector toward the sun, z up, y north, x east, azimuth from north."""
    e, a = math.radians(elev_deg), math.radians(azim_deg)
    return np.array([math.sin(a) * math.cos(e), math.cos(a) * math.cos(e), math.sin(e)])


def receiver_axes(pts: np.ndarray, tower_h: float = 150.0):
    """The tower receiver's centre and two unit vectors in its plane (u
    horizontal, v up the receiver) along which aim offsets are measured: the
    receiver faces the field's centroid."""
    rec = np.array([0.0, 0.0, tower_h])
    centroid = np.array([pts[:, 0].mean(), pts[:, 1].mean(), 0.0])
    n = centroid - rec
    n /= np.linalg.norm(n)
    u = np.array([-n[1], n[0], 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    if v[2] < 0:
        u, v = -u, -v
    return rec, u, v


def aim_heliostat(p: np.ndarray, target: np.ndarray, sv: np.ndarray):
    """Normal of a heliostat at p that reflects the sun (unit vector sv) to
    target -- the bisector of the sun and target directions
    (app/deploy/api/examples/tower-demo.py:49-54) -- and its slant range."""
    to_t = target - p
    slant = float(np.linalg.norm(to_t))
    nrm = to_t / slant + sv
    return nrm / np.linalg.norm(nrm), slant


def tower(n_helio: int = 1000, raycount: int = 10_000_000, helio_size: float = 10.0,
          rec_size: float = 15.0, tower_h: float = 150.0, r_max: float = 1000.0,
          sun_elev: float = 45.0, sun_azim: float = 180.0, csr: float = 0.1, slope: float = 1.5,
          refl: float = 0.9, rec_abs: float = 0.95, sun: str = "buie", dyn_group: bool = True,
          aim_offsets=None) -> Scene:
    """Heliostat field in the north half-plane, flat receiver facing it.
    aim_offsets (n_helio x 2, metres, optional): each heliostat aims at the
    receiver centre plus du * u + dv * v (receiver_axes) instead of the centre;
    its focal length follows the aim point's slant range."""
    sv = sun_vector(sun_elev, sun_azim)
    pts = staggered_layout(n_helio, 100.0, r_max, 2.2 * helio_size, north_only=True)
    rec, ax_u, ax_v = receiver_axes(pts, tower_h)
    s = Scene(sun_xyz=tuple(sv * 1000.0), raycount=raycount, maxcount=2**31 - 1, dyn_group=dyn_group)
    if sun == "buie":
        a, v = buie_profile(csr)
        s.sun_shape, s.sun_sigma, s.sun_table = "d", 4.65, (a, v)
    else:
        s.sun_shape, s.sun_sigma = "p", 4.65
    s.optics = [Optic("helio", OpticFace(dist="g", refl=refl, slope=slope, spec=0.0),
                      OpticFace(dist="g", refl=0.0, slope=slope, spec=0.0)),
                Optic("receiver", OpticFace(dist="g", refl=1.0 - rec_abs), OpticFace(dist="g", refl=1.0 - rec_abs))]
    els = []
    for i, (x, y) in enumerate(pts):
        p = np.array([x, y, 0.0])
        target = rec if aim_offsets is None else rec + aim_offsets[i][0] * ax_u + aim_offsets[i][1] * ax_v
        nrm, slant = aim_heliostat(p, target, sv)
        els.append(Element(xyz=tuple(p), aim=tuple(p + 100.0 * nrm), aperture="r",
                           ap_params=(helio_size, helio_size), surface="s",
                           surf_params=(1.0 / (2.0 * slant),), optic="helio"))
    centroid = np.array([pts[:, 0].mean(), pts[:, 1].mean(), 0.0])
    recel = Element(xyz=tuple(rec), aim=tuple(centroid), aperture="r", ap_params=(rec_size, rec_size),
                    surface="f", optic="receiver")
    s.stages = [Stage(elements=els), Stage(elements=[recel])]
    return s


def facet_zrot(nrm, u) -> float:
    """The z rotation (degrees) that turns an element aimed along unit
    vector `nrm` so its local x axis lies along `u` (in the element's plane).
    With zrot 0 the local x axis is (cos a, 0, -sin a) and y is
    (-sin a sin b, cos b, -cos a sin b), a = atan2(nx, nz), b = asin(ny)
    (coretrace/input.cpp:174-198, mathproc.cpp:175-217); zrot g turns x to
    cos(g) x0 - sin(g) y0."""
    a, b = math.atan2(nrm[0], nrm[2]), math.asin(nrm[1])
    x0 = np.array([math.cos(a), 0.0, -math.sin(a)])
    y0 = np.array([-math.sin(a) * math.sin(b), math.cos(b), -math.cos(a) * math.sin(b)])
    return math.degrees(math.atan2(-float(u @ y0), float(u @ x0)))


def faceted_field(n_helio: int = 10_000, raycount: int = 50_000_000, facet: float = 5.0,
                  rec_radius: float = 8.5, rec_height: float = 20.0, tower_h: float = 200.0,
                  r_max: float = 1500.0, sun_elev: float = 45.0, sun_azim: float = 180.0,
                  sun_sigma: float = 2.5, slope: float = 1.5, refl: float = 0.9,
                  rec_abs: float = 0.95) -> Scene:
    """Surround field of 2x2-facet heliostats and a cylindrical receiver."""
    sv = sun_vector(sun_elev, sun_azim)
    pts = staggered_layout(n_helio, 120.0, r_max, 2.6 * facet * 2, north_only=False)
    rec = np.array([0.0, 0.0, tower_h])
    s = Scene(sun_shape="g", sun_sigma=sun_sigma, sun_xyz=tuple(sv * 1000.0), raycount=raycount,
              maxcount=2**31 - 1, dyn_group=True)
    s.optics = [Optic("helio", OpticFace(dist="g", refl=refl, slope=slope, spec=0.0),
                      OpticFace(dist="g", refl=0.0, slope=slope, spec=0.0)),
                Optic("receiver", OpticFace(dist="g", refl=1.0 - rec_abs), OpticFace(dist="g", refl=1.0 - rec_abs))]
    els = []
    for x, y in pts:
        p = np.array([x, y, 0.0])
        to_rec = rec - p
        nrm = to_rec / np.linalg.norm(to_rec) + sv
        nrm /= np.linalg.norm(nrm)
        # in-plane axes: u horizontal, w = n x u
        u = np.cross([0.0, 0.0, 1.0], nrm)
        u /= np.linalg.norm(u)
        w = np.cross(nrm, u)
        zrot = facet_zrot(nrm, u)
        for du, dw in ((-0.5, -0.5), (0.5, -0.5), (-0.5, 0.5), (0.5, 0.5)):
            c = p + (du * u + dw * w) * (facet + 0.05)
            els.append(Element(xyz=tuple(c), aim=tuple(c + 100.0 * nrm), zrot=zrot, aperture="r",
                               ap_params=(facet, facet), surface="f", optic="helio"))
    recel = Element(xyz=(0.0, -rec_radius, tower_h), aim=(0.0, 1000.0 - rec_radius, tower_h), aperture="l",
                    ap_params=(0.0, 0.0, rec_height), surface="t", surf_params=(1.0 / rec_radius,),
                    optic="receiver")
    s.stages = [Stage(elements=els), Stage(elements=[recel])]
    return s


# -------------------------------------------------------- batched scenes ---


def sun_position(lat: float, day: float, hour: float) -> np.ndarray:
    """Unit vector toward the sun in this module's frame (x east, y north, z
    up) from latitude, day of year and solar hour, by the formula of the
    reference's st_sun_position (coretrace/stapi.cpp:500-527), whose frame is
    x west, y zenith, z north."""
    d2r = math.pi / 180.0
    decl = 180.0 / math.pi * math.asin(0.39795 * math.cos(0.98563 * d2r * (day - 173)))
    ha = 15.0 * (hour - 12.0)
    el = 180.0 / math.pi * math.asin(math.sin(decl * d2r) * math.sin(lat * d2r) +
                                      math.cos(decl * d2r) * math.cos(ha * d2r) * math.cos(lat * d2r))
    az = 180.0 / math.pi * math.acos((math.sin(d2r * decl) * math.cos(d2r * lat) -
                                      math.cos(d2r * decl) * math.sin(d2r * lat) * math.cos(d2r * ha)) /
                                     math.cos(d2r * el) + 0.0000000001)
    if math.sin(ha * d2r) > 0.0:
        az = 360.0 - az
    x = -math.sin(az * d2r) * math.cos(el * d2r)
    y = math.sin(el * d2r)
    z = math.cos(az * d2r) * math.cos(el * d2r)
    return np.array([-x, z, y])


def zernike_mon(path: str, focal: float, x0: float, y0: float, rng: np.random.Generator, order: int = 6,
                sigma: float = 1e-4) -> str:
    """Write a .mon monomial surface file (coretrace/input.cpp:395-414: the
    order, then B_ij for i = 0..order, j = 0..i, z = sum B_ij x^j y^(i-j))
    for the patch of the paraboloid z = r^2 / (4 focal) around (x0, y0), in
    the frame of a facet at (x0, y0, z(x0, y0)) parallel to the axis, with
    every coefficient perturbed by N(0, sigma)."""
    vals = []
    for i in range(order + 1):
        for j in range(i + 1):
            b = rng.normal(0.0, sigma)
            if i == 1:
                b += (x0 if j == 1 else y0) / (2.0 * focal)
            if i == 2 and j in (0, 2):
                b += 1.0 / (4.0 * focal)
            vals.append(b)
    with open(path, "w") as f:
        f.write(f"{order}\n" + "".join(f"{v:.12g}\n" for v in vals))
    return path


def dish(sun: np.ndarray, raycount: int, mon_dir: str, rng: np.random.Generator, focal: float = 4.0,
         sun_shape: str = "buie") -> Scene:
    """A sun-tracking dish of focal length `focal`: seven order-6 Zernike
    facets ('m', 2.2 m circles: one on the axis, six on a ring of 2.2 m, each
    its patch of the paraboloid with coefficients perturbed by N(0, 1e-4)),
    a toroidal rim of four 90-degree torus sections ('d', annular radius
    3.6 m, tube radius 0.25 m, 'a' aperture with A = B = 0, turned by
    0/90/180/270 degrees), and a flat 1.5 x 1.5 m receiver at the focus.  The
    stage frame's z axis points at the sun."""
    s = Scene(sun_xyz=tuple(sun * 1000.0), raycount=raycount, maxcount=2 ** 31 - 1)
    if sun_shape == "buie":
        a, v = buie_profile(0.1)
        s.sun_shape, s.sun_sigma, s.sun_table = "d", 4.65, (a, v)
    else:
        s.sun_shape, s.sun_sigma = "p", 4.65
    s.optics = [Optic("mirror", OpticFace(dist="g", refl=0.93, slope=1.0, spec=0.2),
                      OpticFace(dist="g", refl=0.0, slope=1.0, spec=0.2)),
                Optic("receiver", OpticFace(dist="g", refl=0.05), OpticFace(dist="g", refl=0.05))]
    els = []
    spots = [(0.0, 0.0)] + [(2.2 * math.cos(q * math.pi / 3), 2.2 * math.sin(q * math.pi / 3)) for q in range(6)]
    for q, (x, y) in enumerate(spots):
        z = (x * x + y * y) / (4.0 * focal)
        mon = zernike_mon(os.path.join(mon_dir, f"dish{id(s)}_{q}.mon"), focal, x, y, rng)
        els.append(Element(xyz=(x, y, z), aim=(x, y, z + 1.0), aperture="c", ap_params=(2.2,), surface="m",
                           surf_file=mon, optic="mirror"))
    for q in range(4):
        els.append(Element(xyz=(0.0, 0.0, 0.6), aim=(0.0, 0.0, 1.6), zrot=90.0 * q, aperture="a",
                           ap_params=(0.0, 0.0, 90.0), surface="d", surf_params=(3.6, 0.25), optic="mirror"))
    recv = Element(xyz=(0.0, 0.0, focal), aim=(0.0, 0.0, 0.0), aperture="r", ap_params=(1.5, 1.5), surface="f",
                   optic="receiver")
    s.stages = [Stage(aim=tuple(sun), elements=els), Stage(aim=tuple(sun), elements=[recv])]
    return s


def batch_scenes(n: int = 256, seed: int = 2026, hits=(100_000, 2_000_000), mon_dir: str = ".",
                 helio=(200, 5000), lat: float = 35.0) -> list:
    """The batched-scene workload: n scenes drawn from `seed`, each a tower
    field of U[200, 5000] spherical heliostats (as :func:`tower`, Buie CSR 0.1
    sun) or a dish (:func:`dish`), with equal odds; the sun from
    st_sun_position at latitude 35 N, a uniform day of the year and a solar
    hour with the sun at least 15 degrees up; U[hits] ray hits.  Returns the
    scenes; .mon files are written to mon_dir."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        while True:
            day, hour = float(rng.uniform(1, 366)), float(rng.uniform(6, 18))
            sun = sun_position(lat, day, hour)
            if sun[2] > math.sin(math.radians(15.0)):
                break
        nh = int(rng.integers(hits[0], hits[1] + 1))
        if rng.random() < 0.5:
            nhel = int(rng.integers(helio[0], helio[1] + 1))
            el = math.degrees(math.asin(sun[2]))
            az = math.degrees(math.atan2(sun[0], sun[1])) % 360.0
            sc = tower(n_helio=nhel, raycount=nh, r_max=1000.0 * math.sqrt(nhel / 1000.0), sun_elev=el,
                       sun_azim=az)
        else:
            sc = dish(sun, nh, mon_dir, rng)
        out.append(sc)
    return out
