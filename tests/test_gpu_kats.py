"""The intercept-factor known answers of the reference's trough example
(see test_reference_kats.py) on the HIP library, both precisions: 1 and 0
exactly, 0.7066204592546165 within 3 standard deviations of binomial noise
(the library draws from its own counter-based streams)."""
import math

import pytest

import test_reference_kats as K

pytestmark = pytest.mark.gpu


def _setup(bits):
    def f(lib, ctx):
        lib.dll.st_amd_set_precision(ctx, bits)
    return f


def _if(lib, bits, angle, sun):
    import harness
    orig = harness.run_scene

    def run(l, sc, seed, setup=None, after=None):
        return orig(l, sc, seed, setup=_setup(bits), after=after)
    K.run_scene = run
    try:
        return K._intercept_factor(lib, angle, sun)
    finally:
        K.run_scene = orig


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("nominal,sun_xyz", [(0.0, (0.0, 0.0, 100.0)), (90.0, (100.0, 0.0, 0.0))])
def test_gpu_intercept_factor_kats(hip_lib, bits, nominal, sun_xyz):
    assert _if(hip_lib, bits, nominal + 0.0, sun_xyz) == 1.0
    assert _if(hip_lib, bits, nominal + 2.5, sun_xyz) == 0.0
    got = _if(hip_lib, bits, nominal + K.ERR_RAD_DEG, sun_xyz)
    sigma = math.sqrt(K.KAT * (1 - K.KAT) / 985989)
    assert abs(got - K.KAT) < 3 * sigma, (got, K.KAT, sigma)
