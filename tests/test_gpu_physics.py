"""The physics checks of tests/physics_checks.py on the HIP library, FP64
and FP32 paths (positions and directions of the FP32 path are held to 1e-5
m and 1e-8 instead of 1e-9 m and 1e-12)."""
import pytest

import physics_checks as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("check", P.ALL, ids=[c.__name__ for c in P.ALL])
def test_gpu_physics(hip_lib, check, bits):
    def setup(lib, ctx):
        lib.dll.st_amd_set_precision(ctx, bits)
    check(hip_lib, setup=setup, tol=1.0 if bits == 64 else 1e4)
