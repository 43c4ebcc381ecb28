import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from soltrace_amd import stapi  # noqa: E402

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "st_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "coretrace_ref.so")
HIP_SO = stapi.DEFAULT_LIB


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    if not os.path.exists(ORACLE_SO):
        pytest.skip("oracle not built (make -C oracle)")
    lib = stapi.StLib(ORACLE_SO)
    import ctypes as C
    lib.dll.st_oracle_set_rng.argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_oracle_set_rng.restype = C.c_int
    lib.dll.st_oracle_set_torus_start.argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_oracle_set_torus_start.restype = C.c_int
    lib.dll.st_oracle_ray_sun_index.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    lib.dll.st_oracle_ray_sun_index.restype = C.c_int
    lib.dll.st_oracle_flux.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.dll.st_oracle_flux.restype = C.c_int
    return lib


@pytest.fixture(scope="session")
def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("reference build absent (make -C oracle ref needs /root/reference)")
    return stapi.StLib(REF_SO)


@pytest.fixture(scope="session")
def hip_lib():
    if not gpu_available():
        pytest.skip("no GPU")
    lib = stapi.load()
    import ctypes as C
    lib.dll.st_amd_set_precision.argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_amd_set_cull.argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_amd_flux_request.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_double, C.c_double, C.c_double, C.c_double]
    lib.dll.st_amd_flux_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_double)]
    lib.dll.st_amd_flux_by_source.argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_amd_flux_source_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.c_longlong]
    lib.dll.st_amd_flux_source_result.restype = C.c_longlong
    lib.dll.st_amd_ray_sun_index.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), C.c_int]
    lib.dll.st_amd_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
    return lib
