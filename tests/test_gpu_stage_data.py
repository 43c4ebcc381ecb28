"""Stage-data save/resume on the GPU against the oracle in counter mode:
records, the saved stage-1 rows and the saved stage-2 array (FP64: stage,
element and ray numbers exact, positions and cosines within 1e-6), a run
resumed from saved rows, and the reference's C++ entry point
st_sim_run_data called from a C++ program."""
import os
import subprocess

import numpy as np
import pytest

from harness import assert_records_match, counter_mode
from test_stage_data import run_data, three_stage_scene

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rows_match(a, b):
    assert a.shape == b.shape, (a.shape, b.shape)
    np.testing.assert_array_equal(a[:, 6:], b[:, 6:])
    assert np.all(np.abs(a[:, :6] - b[:, :6]) <= 1e-6 + 1e-6 * np.abs(b[:, :6])), np.abs(a[:, :6] - b[:, :6]).max()


@pytest.mark.parametrize("seed,n_stages", [(3, 3), (12, 3), (3, 2)])
def test_save_hip_equals_oracle(hip_lib, oracle_lib, seed, n_stages):
    # n_stages 2: stage 2 is the last stage and still hands its rays on to
    # the saved stage-2 array
    sc = three_stage_scene(4000, n_stages=n_stages)
    print(f"save seed {seed}, {n_stages} stages", flush=True)
    o = run_data(oracle_lib, sc, seed, save=True, setup=counter_mode)
    g = run_data(hip_lib, sc, seed, save=True)
    assert g[0] == o[0] > 0, (g[0], o[0], g[3])
    assert_records_match(g[1], o[1], pos_tol=1e-6, cos_tol=1e-6)
    rows_match(g[2][0], o[2][0])
    rows_match(g[2][1], o[2][1])
    assert len(o[2][1]) > 0


def test_load_hip_equals_oracle(hip_lib, oracle_lib):
    sc = three_stage_scene(4000)
    _, _, (s1, s2), _ = run_data(oracle_lib, sc, 5, save=True, setup=counter_mode)
    print(f"load {len(s1)} + {len(s2)} rows", flush=True)
    o = run_data(oracle_lib, sc, 21, s1, s2, setup=counter_mode)
    g = run_data(hip_lib, sc, 21, s1, s2)
    assert g[0] == o[0] > len(s1), (g[0], o[0], g[3])
    assert_records_match(g[1], o[1], pos_tol=1e-6, cos_tol=1e-6)
    assert g[1].nsunrays == int(s1[-1, 7])


def test_load_failure_matches_oracle(hip_lib, oracle_lib):
    sc = three_stage_scene(500)
    _, _, (s1, s2), _ = run_data(oracle_lib, sc, 5, save=True, setup=counter_mode)
    s2 = s2.copy()
    s2[-1, 6] = 0.0
    o = run_data(oracle_lib, sc, 9, s1, s2, setup=counter_mode)
    g = run_data(hip_lib, sc, 9, s1, s2)
    assert g[0] == o[0] == -1 and g[3][-1] == o[3][-1]


def test_resumed_flux_map_equals_saving_run(hip_lib):
    # a flux map on a stage-1 facet: binned at record time while tracing,
    # from the loaded rows when resuming; the same records give the same map
    import ctypes as C
    from soltrace_amd import stapi
    d = hip_lib.dll
    d.st_amd_flux_request.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_double]
    d.st_amd_flux_result.argtypes = [C.c_void_p, C.POINTER(C.c_uint), C.POINTER(C.c_double)]
    sc = three_stage_scene(4000, flat=True)

    def run(s1=None, s2=None, save=False):
        ctx = hip_lib.st_create_context()
        try:
            sc.apply(hip_lib, ctx)
            assert d.st_amd_flux_request(ctx, 0, 4, 20, 20, 0, -0.5, 0.5, -0.5, 0.5) == 1
            n = stapi.run_data(hip_lib, ctx, 8, s1, s2, save)
            buf = (C.c_uint * 400)()
            assert d.st_amd_flux_result(ctx, buf, None) == 400
            return n, np.frombuffer(buf, dtype=np.uint32).copy(), stapi.saved_stage_data(hip_lib, ctx)
        finally:
            hip_lib.st_free_context(ctx)

    print("flux map: save, then resume", flush=True)
    n, f_saved, (s1, s2) = run(save=True)
    m, f_resumed, _ = run(s1, s2)
    assert n > 0 and m > 0 and f_saved.sum() > 0
    np.testing.assert_array_equal(f_resumed, f_saved)


CXX = r"""
#include <cstdio>
#include <vector>
#include "stapi.h"
#include "stapi_amd.h"
// save, then resume from the saved vectors, through the reference's C++
// entry point; print the counts and a checksum of each
int main(int argc, char **argv) {
    st_context_t c = st_create_context();
    if (st_amd_read_stinput(c, argv[1]) != 1) return 2;
    st_sim_params(c, 3000, 1000000, 0);
    st_sim_errors(c, 1, 1);
    std::vector<std::vector<double> > s1, s2;
    int n = st_sim_run_data(c, 4, &s1, &s2, true, 0, 0);
    double k1 = 0, k2 = 0;
    for (auto &r : s1) for (double v : r) k1 += v;
    for (auto &r : s2) for (double v : r) k2 += v;
    printf("%d %zu %zu %.17g %.17g\n", n, s1.size(), s2.size(), k1, k2);
    int m = st_sim_run_data(c, 6, &s1, &s2, false, 0, 0);
    printf("%d %d\n", m, st_num_intersections(c));
    st_free_context(c);
    return n < 0 || m < 0;
}
"""


def test_cpp_entry_point_saves_and_resumes(hip_lib, tmp_path):
    import shutil
    from soltrace_amd import stapi
    from test_strace import GOLDEN
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src, exe = tmp_path / "app.cpp", tmp_path / "app"
    src.write_text(CXX)
    so = os.path.join(ROOT, "soltrace_amd", "coretrace_api.so")
    subprocess.run(["g++", "-O1", "-I", os.path.join(ROOT, "include"), str(src), so,
                    "-Wl,-rpath," + os.path.dirname(so), "-o", str(exe)], check=True)
    stinput = os.path.join(GOLDEN, "lens.stinput")
    print("c++ st_sim_run_data on lens", flush=True)
    r = subprocess.run([str(exe), stinput], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    l1, l2 = r.stdout.split("\n")[:2]
    n, n1, n2, k1, k2 = l1.split()
    # the same through the plain-array entry points
    import ctypes as C
    d = hip_lib.dll
    d.st_amd_read_stinput.argtypes = [C.c_void_p, C.c_char_p]
    ctx = hip_lib.st_create_context()
    try:
        assert d.st_amd_read_stinput(ctx, stinput.encode()) == 1
        hip_lib.st_sim_params(ctx, 3000, 1000000, 0)
        hip_lib.st_sim_errors(ctx, 1, 1)
        assert stapi.run_data(hip_lib, ctx, 4, save=True) == int(n)
        a1, a2 = stapi.saved_stage_data(hip_lib, ctx)
        assert (len(a1), len(a2)) == (int(n1), int(n2)) and len(a2) > 0
        # the same rows: their sums agree to the summation order
        assert abs(float(k1) - a1.sum()) <= 1e-9 * abs(a1).sum()
        assert abs(float(k2) - a2.sum()) <= 1e-9 * abs(a2).sum()
        m = stapi.run_data(hip_lib, ctx, 6, a1, a2)
    finally:
        hip_lib.st_free_context(ctx)
    assert l2.split() == [str(m), str(m)]
