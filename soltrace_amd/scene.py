"""Scene description that can be replayed into any stapi library.

A :class:`Scene` holds what a SolTrace user sets through the C API: sun,
optics, stages, elements and simulation parameters.  :meth:`Scene.apply`
issues the st_* calls in the order the reference's callers use
(app/deploy/api/pysoltrace.py:1399-1408, 735-776: the aperture is set before
the surface parameters, coretrace/input.cpp:295-298).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from .stapi import StLib


@dataclass
class OpticFace:
    dist: str = "g"
    optnum: int = 1
    apgr: int = 0
    order: int = 0
    rreal: float = 1.0
    rimag: float = 0.0
    refl: float = 0.0
    trans: float = 0.0
    grating: Sequence[float] = (0.0, 0.0, 0.0, 0.0)
    slope: float = 0.0      # mrad
    spec: float = 0.0       # mrad
    refl_table: Optional[Sequence[tuple]] = None   # [(angle_mrad, value), ...]
    trans_table: Optional[Sequence[tuple]] = None


@dataclass
class Optic:
    name: str
    front: OpticFace = field(default_factory=OpticFace)
    back: OpticFace = field(default_factory=OpticFace)


@dataclass
class Element:
    xyz: Sequence[float] = (0.0, 0.0, 0.0)
    aim: Sequence[float] = (0.0, 0.0, 1.0)
    zrot: float = 0.0
    aperture: str = "r"
    ap_params: Sequence[float] = (1.0, 1.0, 0, 0, 0, 0, 0, 0)
    surface: str = "f"
    surf_params: Sequence[float] = (0, 0, 0, 0, 0, 0, 0, 0)
    surf_file: Optional[str] = None
    interaction: int = 2
    optic: str = "opt"
    enabled: bool = True


@dataclass
class Stage:
    xyz: Sequence[float] = (0.0, 0.0, 0.0)
    aim: Sequence[float] = (0.0, 0.0, 1.0)
    zrot: float = 0.0
    virtual: bool = False
    multihit: bool = True
    tracethrough: bool = False
    elements: list = field(default_factory=list)


@dataclass
class Scene:
    sun_shape: str = "p"
    sun_sigma: float = 4.65
    sun_xyz: Sequence[float] = (0.0, 0.0, 100.0)
    point_source: bool = False
    sun_table: Optional[tuple] = None   # (angles_mrad, intensities)
    optics: list = field(default_factory=list)
    stages: list = field(default_factory=list)
    raycount: int = 10000
    maxcount: int = 10_000_000
    dyn_group: bool = False
    sunshape_err: bool = True
    optics_err: bool = True

    def n_elements(self) -> int:
        return sum(len(s.elements) for s in self.stages)

    def apply(self, lib: StLib, ctx) -> None:
        """Replay this scene into context ``ctx`` of library ``lib``."""
        def ok(rc, what):
            if rc is not None and rc < 0:
                raise RuntimeError(f"{what} failed ({rc})")
        ok(lib.st_sun(ctx, int(self.point_source), self.sun_shape.encode(), float(self.sun_sigma)), "st_sun")
        ok(lib.st_sun_xyz(ctx, *map(float, self.sun_xyz)), "st_sun_xyz")
        if self.sun_table is not None:
            a = np.ascontiguousarray(self.sun_table[0], dtype=np.float64)
            v = np.ascontiguousarray(self.sun_table[1], dtype=np.float64)
            ok(lib.st_sun_userdata(ctx, len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
                                   v.ctypes.data_as(C.POINTER(C.c_double))), "st_sun_userdata")
        for o in self.optics:
            idx = lib.st_add_optic(ctx, o.name.encode())
            for fb, face in ((1, o.front), (2, o.back)):
                g = (C.c_double * 4)(*face.grating)
                rt = face.refl_table or []
                tt = face.trans_table or []
                ra = (C.c_double * max(1, len(rt)))(*[p[0] for p in rt])
                rv = (C.c_double * max(1, len(rt)))(*[p[1] for p in rt])
                ta = (C.c_double * max(1, len(tt)))(*[p[0] for p in tt])
                tv = (C.c_double * max(1, len(tt)))(*[p[1] for p in tt])
                ok(lib.st_optic(ctx, idx, fb, face.dist.encode(), face.optnum, face.apgr, face.order,
                                face.rreal, face.rimag, face.refl, face.trans, g, face.slope, face.spec,
                                int(bool(rt)), len(rt), ra, rv, int(bool(tt)), len(tt), ta, tv), "st_optic")
        for si, st in enumerate(self.stages):
            ok(lib.st_add_stage(ctx), "st_add_stage")
            ok(lib.st_stage_xyz(ctx, si, *map(float, st.xyz)), "st_stage_xyz")
            ok(lib.st_stage_aim(ctx, si, *map(float, st.aim)), "st_stage_aim")
            ok(lib.st_stage_zrot(ctx, si, float(st.zrot)), "st_stage_zrot")
            ok(lib.st_stage_flags(ctx, si, int(st.virtual), int(st.multihit), int(st.tracethrough)),
               "st_stage_flags")
            if st.elements:
                ok(lib.st_add_elements(ctx, si, len(st.elements)), "st_add_elements")
            for ei, el in enumerate(st.elements):
                ok(lib.st_element_enabled(ctx, si, ei, int(el.enabled)), "enabled")
                ok(lib.st_element_xyz(ctx, si, ei, *map(float, el.xyz)), "xyz")
                ok(lib.st_element_aim(ctx, si, ei, *map(float, el.aim)), "aim")
                ok(lib.st_element_zrot(ctx, si, ei, float(el.zrot)), "zrot")
                ok(lib.st_element_aperture(ctx, si, ei, el.aperture.encode()), "aperture")
                ap = (C.c_double * 8)(*[float(v) for v in list(el.ap_params) + [0] * (8 - len(el.ap_params))])
                ok(lib.st_element_aperture_params(ctx, si, ei, ap), "aperture_params")
                ok(lib.st_element_surface(ctx, si, ei, el.surface.encode()), "surface")
                if el.surf_file:
                    ok(lib.st_element_surface_file(ctx, si, ei, el.surf_file.encode()), "surface_file")
                else:
                    sp = (C.c_double * 8)(*[float(v) for v in
                                            list(el.surf_params) + [0] * (8 - len(el.surf_params))])
                    ok(lib.st_element_surface_params(ctx, si, ei, sp), "surface_params")
                ok(lib.st_element_interaction(ctx, si, ei, el.interaction), "interaction")
                ok(lib.st_element_optic(ctx, si, ei, el.optic.encode()), "optic")
        ok(lib.st_sim_params(ctx, int(self.raycount), int(self.maxcount), int(self.dyn_group)), "st_sim_params")
        ok(lib.st_sim_errors(ctx, int(self.sunshape_err), int(self.optics_err)), "st_sim_errors")
