"""Calls of the element-statistics entry point, shared by the CPU and GPU
tests: st_amd_element_stats (library) and st_oracle_element_stats (oracle)
take the same arguments and fill the same stats[15]
(app/src/project.cpp:1139-1436)."""
import ctypes as C

import numpy as np

NAMES = ["power_per_ray", "peak", "min", "ave", "sigma", "uniformity", "peak_unc", "ave_unc",
         "cx", "cy", "cz", "binned", "zval", "bx", "by"]


def element_stats(lib, ctx, fn, stage, elem, nx, ny, autoscale, absorbed_only, dni, bounds=(0.0, 0.0, 0.0, 0.0)):
    f = getattr(lib.dll, fn)
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                  C.POINTER(C.c_double), C.POINTER(C.c_uint), C.POINTER(C.c_double)]
    f.restype = C.c_int
    b = (C.c_double * 4)(*bounds)
    grid = (C.c_uint * (nx * ny))()
    st = (C.c_double * 15)()
    rc = f(ctx, stage, elem, nx, ny, autoscale, absorbed_only, dni, b, grid, st)
    if rc != 1:
        return None
    return (np.array(list(b)), np.frombuffer(grid, dtype=np.uint32).reshape(nx, ny).copy(),
            dict(zip(NAMES, list(st))))
