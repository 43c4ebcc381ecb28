"""The operation-counting build of the oracle (oracle/st_count.h,
st_oracle_count.cpp), which bench.py uses for the algorithmic flops behind
each kernel's roofline figure:
  * it runs the same arithmetic: its records equal the oracle's bit for bit;
  * it counts what the operations are: st_transform_to_local is 3
    subtractions and two 3x3 matrix-vector products, 15 additions and 18
    multiplications (coretrace/mathproc.cpp:58-72, 115-142);
  * its phases add up: per-phase counts sum to the whole run's."""
import ctypes as C
import os

import numpy as np
import pytest

from harness import counter_mode, run_scene
from soltrace_amd import stapi, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COUNT_SO = os.path.join(ROOT, "oracle", "_build", "st_oracle_count.so")


@pytest.fixture(scope="module")
def count_lib():
    if not os.path.exists(COUNT_SO):
        pytest.skip("make -C oracle")
    lib = stapi.StLib(COUNT_SO)
    for n in ("st_oracle_set_rng", "st_oracle_set_torus_start"):
        getattr(lib.dll, n).argtypes = [C.c_void_p, C.c_int]
    lib.dll.st_opcount_read.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
    return lib


def read(lib, tests=False):
    out = (C.c_ulonglong * 83)()
    assert lib.dll.st_opcount_read(out, 83) == 83
    a = np.frombuffer(out, dtype=np.uint64).astype(np.int64)
    return (a[:80].reshape(16, 5), a[80:]) if tests else a[:80].reshape(16, 5)


@pytest.mark.parametrize("name", ["trough", "tower"])
def test_same_records_as_the_oracle(count_lib, oracle_lib, name):
    sc = workloads.trough(raycount=3000) if name == "trough" else workloads.tower(n_helio=40, raycount=2000,
                                                                                   r_max=300.0)
    for setup in (None, counter_mode):
        a = run_scene(count_lib, sc, 5, setup=setup)
        b = run_scene(oracle_lib, sc, 5, setup=setup)
        assert a[0] == b[0] > 0
        for f in ("pos", "cos"):
            assert np.array_equal(getattr(a[1], f).view(np.uint64), getattr(b[1], f).view(np.uint64))
        for f in ("element", "stage", "raynum"):
            np.testing.assert_array_equal(getattr(a[1], f), getattr(b[1], f))


def test_transform_counts_its_33_flops(count_lib):
    count_lib.dll.st_opcount_reset()
    D3, M33 = C.c_double * 3, (C.c_double * 3) * 3
    R = M33(D3(1, 2, 3), D3(4, 5, 6), D3(7, 8, 9))
    pl, cl = D3(), D3()
    count_lib.st_transform_to_local(D3(1, 1, 1), D3(0, 0, 1), D3(0.5, 0, 0), R, pl, cl)
    n = read(count_lib)
    assert n[2].tolist() == [15, 18, 0, 0, 0]
    assert n.sum() == 33
    assert list(pl) == [0.5 + 2 + 3, 2 + 5 + 6, 3.5 + 8 + 9]  # R (p - origin)


def test_phases_partition_the_run(count_lib):
    sc = workloads.tower(n_helio=40, raycount=2000, r_max=300.0)
    count_lib.dll.st_opcount_reset()
    _, res, _, _ = run_scene(count_lib, sc, 5, setup=counter_mode)
    n, tests = read(count_lib, tests=True)
    assert n[15].sum() == 0  # every sun ray's search was settled
    for p in (0, 1, 2, 3, 5, 12, 13):
        assert n[p].sum() > 0, p
    # brute force: every sun ray tests the 40 heliostats at least once; every
    # ray entering stage 2 tests the receiver, again after a reflection
    n2 = (res.stage == 2).sum()
    assert tests[0] >= 40 * res.nsunrays and n2 <= tests[1] <= n2 + (res.element[res.stage == 2] > 0).sum()
