"""ctypes binding of SolTrace's C API (coretrace/stapi.h:76-171).

The binding is generic: it declares the 55 stapi entry points on any shared
library that exports them.  :func:`load` opens the HIP build of this package
(``soltrace_amd/coretrace_api.so``); tests open the checkers the same way.

Argument and return conventions follow the reference header exactly
(``st_uint_t`` is ``unsigned long``, 64-bit on LP64, coretrace/stapi.h:72).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

st_uint_t = C.c_ulong
st_context_t = C.c_void_p
CALLBACK = C.CFUNCTYPE(C.c_int, st_uint_t, st_uint_t, st_uint_t, st_uint_t, st_uint_t, C.c_void_p)

_D3 = C.c_double * 3
_M33 = (C.c_double * 3) * 3
_D8 = C.c_double * 8

_SIGS = {
    "st_create_context": (st_context_t, []),
    "st_free_context": (C.c_int, [st_context_t]),
    "st_num_messages": (C.c_int, [st_context_t]),
    "st_message": (C.c_char_p, [st_context_t, st_uint_t]),
    "st_clear_messages": (None, [st_context_t]),
    "st_num_optics": (C.c_int, [st_context_t]),
    "st_add_optic": (C.c_int, [st_context_t, C.c_char_p]),
    "st_delete_optic": (C.c_int, [st_context_t, st_uint_t]),
    "st_clear_optics": (C.c_int, [st_context_t]),
    "st_optic": (C.c_int, [st_context_t, st_uint_t, C.c_int, C.c_char, C.c_int, C.c_int, C.c_int,
                           C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                           C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double),
                           C.POINTER(C.c_double)]),
    "st_num_stages": (C.c_int, [st_context_t]),
    "st_add_stage": (C.c_int, [st_context_t]),
    "st_add_stages": (C.c_int, [st_context_t, st_uint_t]),
    "st_delete_stage": (C.c_int, [st_context_t, st_uint_t]),
    "st_clear_stages": (C.c_int, [st_context_t]),
    "st_stage_xyz": (C.c_int, [st_context_t, st_uint_t, C.c_double, C.c_double, C.c_double]),
    "st_stage_aim": (C.c_int, [st_context_t, st_uint_t, C.c_double, C.c_double, C.c_double]),
    "st_stage_zrot": (C.c_int, [st_context_t, st_uint_t, C.c_double]),
    "st_stage_flags": (C.c_int, [st_context_t, st_uint_t, C.c_int, C.c_int, C.c_int]),
    "st_num_elements": (C.c_int, [st_context_t, st_uint_t]),
    "st_add_element": (C.c_int, [st_context_t, st_uint_t]),
    "st_add_elements": (C.c_int, [st_context_t, st_uint_t, st_uint_t]),
    "st_delete_element": (C.c_int, [st_context_t, st_uint_t, st_uint_t]),
    "st_clear_elements": (C.c_int, [st_context_t, st_uint_t]),
    "st_element_enabled": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_int]),
    "st_element_xyz": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_double, C.c_double, C.c_double]),
    "st_element_aim": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_double, C.c_double, C.c_double]),
    "st_element_zrot": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_double]),
    "st_element_aperture": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_char]),
    "st_element_aperture_params": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.POINTER(C.c_double)]),
    "st_element_surface": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_char]),
    "st_element_surface_params": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.POINTER(C.c_double)]),
    "st_element_surface_file": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_char_p]),
    "st_element_interaction": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_int]),
    "st_element_optic": (C.c_int, [st_context_t, st_uint_t, st_uint_t, C.c_char_p]),
    "st_sun": (C.c_int, [st_context_t, C.c_int, C.c_char, C.c_double]),
    "st_sun_xyz": (C.c_int, [st_context_t, C.c_double, C.c_double, C.c_double]),
    "st_sun_position": (C.c_int, [st_context_t, C.c_double, C.c_double, C.c_double,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_sun_userdata": (C.c_int, [st_context_t, st_uint_t, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_num_intersections": (C.c_int, [st_context_t]),
    "st_locations": (C.c_int, [st_context_t, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_cosines": (C.c_int, [st_context_t, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_elementmap": (C.c_int, [st_context_t, C.POINTER(C.c_int)]),
    "st_stagemap": (C.c_int, [st_context_t, C.POINTER(C.c_int)]),
    "st_raynumbers": (C.c_int, [st_context_t, C.POINTER(C.c_int)]),
    "st_sun_stats": (C.c_int, [st_context_t, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "st_sim_params": (C.c_int, [st_context_t, C.c_int, C.c_int, C.c_int]),
    "st_sim_errors": (C.c_int, [st_context_t, C.c_int, C.c_int]),
    "st_sim_run": (C.c_int, [st_context_t, C.c_uint, CALLBACK, C.c_void_p]),
    "st_calc_euler_angles": (None, [_D3, _D3, C.c_double, _D3]),
    "st_transform_to_local": (None, [_D3, _D3, _D3, _M33, _D3, _D3]),
    "st_transform_to_reference": (None, [_D3, _D3, _D3, _M33, _D3, _D3]),
    "st_matrix_vector_mult": (None, [_M33, _D3, _D3]),
    "st_calc_transform_matrices": (None, [_D3, _M33, _M33]),
    "st_matrix_transpose": (None, [_M33, _M33]),
}

STAPI_SYMBOLS = tuple(_SIGS)

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "coretrace_api.so")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class StLib:
    """A loaded library exporting the stapi.h entry points."""

    def __init__(self, path: str, mode: int = 0):
        self.path = path
        self.dll = C.CDLL(path, mode=mode)
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.dll, name)
            fn.restype = res
            fn.argtypes = args

    def __getattr__(self, name):
        return getattr(self.dll, name)


_loaded: dict[str, StLib] = {}


def load(path: str | None = None) -> StLib:
    """Open (once) the HIP build of the stapi library, or another path."""
    path = os.path.abspath(path or DEFAULT_LIB)
    if path not in _loaded:
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950) first")
        _loaded[path] = StLib(path)
    return _loaded[path]


_NOCB = CALLBACK(0)  # NULL function pointer


@dataclass
class Results:
    pos: np.ndarray        # (n, 3) stage-frame hit points
    cos: np.ndarray        # (n, 3) stage-frame incoming directions
    element: np.ndarray    # (n,) 1-based element, negative = absorbed, 0 = miss
    stage: np.ndarray      # (n,) 1-based stage
    raynum: np.ndarray     # (n,) ray number
    sun_box: tuple         # (xmin, xmax, ymin, ymax)
    nsunrays: int

    def __len__(self):
        return int(self.element.shape[0])


def fetch_results(lib: StLib, ctx) -> Results:
    """Copy the records of the last st_sim_run into numpy arrays
    (the getters of coretrace/stapi.cpp:564-659)."""
    n = lib.st_num_intersections(ctx)
    if n < 0:
        raise RuntimeError("st_num_intersections failed")
    x, y, z = (np.zeros(n) for _ in range(3))
    cx, cy, cz = (np.zeros(n) for _ in range(3))
    el, stg, rn = (np.zeros(n, dtype=np.int32) for _ in range(3))
    if n:
        for rc in (lib.st_locations(ctx, _dp(x), _dp(y), _dp(z)),
                   lib.st_cosines(ctx, _dp(cx), _dp(cy), _dp(cz)),
                   lib.st_elementmap(ctx, _ip(el)), lib.st_stagemap(ctx, _ip(stg)),
                   lib.st_raynumbers(ctx, _ip(rn))):
            if rc != 1:
                raise RuntimeError(f"result getter failed ({rc})")
    b = [C.c_double() for _ in range(4)]
    ns = C.c_int()
    lib.st_sun_stats(ctx, *[C.byref(v) for v in b], C.byref(ns))
    return Results(np.stack([x, y, z], 1), np.stack([cx, cy, cz], 1), el, stg, rn,
                   tuple(v.value for v in b), ns.value)


def messages(lib: StLib, ctx) -> list[str]:
    return [lib.st_message(ctx, i).decode() for i in range(lib.st_num_messages(ctx))]


def run(lib: StLib, ctx, seed: int, callback=None) -> int:
    cb = _NOCB if callback is None else CALLBACK(callback)
    return lib.st_sim_run(ctx, seed, cb, None)


def run_data(lib: StLib, ctx, seed: int, s1=None, s2=None, save: bool = False, api=None, callback=None) -> int:
    """st_amd_sim_run_data: a run that loads stage-1 records (rows of 8) and
    stage-2 rays (rows of 7), and/or saves them (the reference's
    st_sim_run_data, coretrace/stapi.cpp:680-728).  `api` is the library
    exporting the call when it is not `lib` itself."""
    d = api if api is not None else lib.dll
    f = d.st_amd_sim_run_data
    f.restype = C.c_int
    f.argtypes = [st_context_t, C.c_uint, C.POINTER(C.c_double), st_uint_t, C.POINTER(C.c_double), st_uint_t,
                  C.c_int, CALLBACK, C.c_void_p]
    a1 = np.ascontiguousarray(np.zeros((0, 8)) if s1 is None else s1, dtype=np.float64).reshape(-1, 8)
    a2 = np.ascontiguousarray(np.zeros((0, 7)) if s2 is None else s2, dtype=np.float64).reshape(-1, 7)
    cb = _NOCB if callback is None else CALLBACK(callback)
    return f(ctx, seed, _dp(a1), a1.shape[0], _dp(a2), a2.shape[0], int(save), cb, None)


def saved_stage_data(lib: StLib, ctx, api=None):
    """Rows the last run_data(save=True) saved: (stage-1 records (n, 8),
    stage-2 ray array (m, 7))."""
    d = api if api is not None else lib.dll
    f = d.st_amd_saved_stage_data
    f.restype = C.c_int
    f.argtypes = [st_context_t, C.c_int, C.POINTER(C.c_double)]
    out = []
    for which, w in ((1, 8), (2, 7)):
        n = f(ctx, which, None)
        if n < 0:
            raise RuntimeError("st_amd_saved_stage_data failed")
        a = np.zeros((n, w))
        if n:
            f(ctx, which, _dp(a))
        out.append(a)
    return tuple(out)


def run_batch(lib: StLib, ctxs, seeds) -> tuple[int, list[int]]:
    """st_amd_sim_run_batch: trace the contexts' scenes together, context k
    with seeds[k].  Returns (return code, per-context record counts or -1)."""
    f = lib.dll.st_amd_sim_run_batch
    f.restype = C.c_int
    f.argtypes = [C.POINTER(st_context_t), C.c_int, C.POINTER(C.c_uint), C.POINTER(C.c_int)]
    n = len(ctxs)
    arr = (st_context_t * n)(*ctxs)
    sd = (C.c_uint * n)(*[int(s) & 0xFFFFFFFF for s in seeds])
    res = (C.c_int * n)()
    rc = f(arr, n, sd, res)
    return rc, list(res)
