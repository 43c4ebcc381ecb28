"""The strace command line on the HIP library (soltrace_amd/strace/strace)
against the same reader and writer on the oracle in counter mode
(oracle/_build/strace_oracle --counter), over the sample systems of
tests/golden/stinput.  Both .rays files are parsed: sun box, sun-ray count,
record count, stage, element and ray number must be equal; positions and
direction cosines within 1e-6 (FP64, the library's default) plus the
printed precision of %.6e."""
import os

import numpy as np
import pytest

from test_strace import FE_FILES, FE_GOLDEN, SAMPLES, STRACE, STRACE_ORACLE, run_cli

pytestmark = pytest.mark.gpu


def parse(text):
    lines = text.splitlines()
    head = np.array([float(v) for v in lines[1].split()])
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[3:]]).reshape(-1, 9)
    return head, rows


@pytest.mark.skipif(not os.path.exists(STRACE_ORACLE), reason="oracle not built (make -C oracle)")
@pytest.mark.parametrize("name", SAMPLES)
def test_strace_hip_equals_oracle(hip_lib, tmp_path, name):
    args = ["3000", "1000000", "21", "1", "1", "1" if name.startswith("Power") else "0"]
    a, b = tmp_path / "hip", tmp_path / "oracle"
    a.mkdir()
    b.mkdir()
    print(f"strace {name} {' '.join(args)}", flush=True)
    got, r1 = run_cli(STRACE, a, name, args)
    want, r2 = run_cli(STRACE_ORACLE, b, name, args, prefix=["--counter"])
    assert r1.returncode == 0 and got is not None, r1.stdout + r1.stderr
    assert r2.returncode == 0 and want is not None, r2.stdout + r2.stderr
    hg, rg = parse(got)
    ho, ro = parse(want)
    assert hg[4] == ho[4] and hg[5] == ho[5] and hg[5] > 0, (hg, ho)
    np.testing.assert_allclose(hg[:4], ho[:4], rtol=1e-6)
    np.testing.assert_array_equal(rg[:, 6:], ro[:, 6:])
    tol = 1e-6 + 1e-6 * np.abs(ro[:, :6])
    assert np.all(np.abs(rg[:, :6] - ro[:, :6]) <= tol), np.abs(rg[:, :6] - ro[:, :6]).max()


@pytest.mark.skipif(not os.path.exists(STRACE_ORACLE), reason="oracle not built (make -C oracle)")
def test_strace_finite_element_hip_equals_oracle(hip_lib, tmp_path):
    # kriging over two meshed dish facets (tests/golden/make_fe_golden.py)
    args = ["2000", "1000000", "13", "1", "1", "0"]
    a, b = tmp_path / "hip", tmp_path / "oracle"
    a.mkdir()
    b.mkdir()
    print("strace fe_dish " + " ".join(args), flush=True)
    got, r1 = run_cli(STRACE, a, "fe_dish", args, src_dir=FE_GOLDEN, extra=FE_FILES)
    want, r2 = run_cli(STRACE_ORACLE, b, "fe_dish", args, prefix=["--counter"], src_dir=FE_GOLDEN, extra=FE_FILES)
    assert r1.returncode == 0 and got is not None, r1.stdout + r1.stderr
    assert r2.returncode == 0 and want is not None, r2.stdout + r2.stderr
    hg, rg = parse(got)
    ho, ro = parse(want)
    assert hg[4] == ho[4] and hg[5] == ho[5] and hg[5] > 0, (hg, ho)
    np.testing.assert_array_equal(rg[:, 6:], ro[:, 6:])
    tol = 1e-6 + 1e-6 * np.abs(ro[:, :6])
    assert np.all(np.abs(rg[:, :6] - ro[:, :6]) <= tol), np.abs(rg[:, :6] - ro[:, :6]).max()


def test_strace_finite_element_hole_fails_the_run(hip_lib, tmp_path):
    # the interpolation finds no zone in the mesh's hole: the run fails with
    # the reference's message and writes no .rays file
    print("strace fe_holed", flush=True)
    got, r = run_cli(STRACE, tmp_path, "fe_holed", ["300", "200000", "5", "1", "1", "0"], src_dir=FE_GOLDEN,
                     extra=FE_FILES)
    assert got is None, r.stdout
    assert r.stdout.splitlines()[-1] == "error: trace error: Interpolation failed to find nearby nodes to ray hit " \
                                        "in FEA surface."


def test_read_stinput_and_ray_files_through_the_library(hip_lib, tmp_path):
    # st_amd_read_stinput + st_sim_run + st_amd_write_ray_file give the CLI's
    # .rays text byte for byte, and the binary file holds the getters' doubles
    import ctypes as C
    import shutil

    from soltrace_amd import stapi
    from test_strace import GOLDEN, read_binary_rays
    d = hip_lib.dll
    d.st_amd_read_stinput.argtypes = [C.c_void_p, C.c_char_p]
    d.st_amd_write_ray_file.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
    args = ["2000", "100000", "5", "1", "1", "0"]
    (tmp_path / "cli").mkdir()
    cli_text, r = run_cli(STRACE, tmp_path / "cli", "lens", args)
    assert r.returncode == 0 and cli_text is not None, r.stdout
    src = tmp_path / "lens.stinput"
    shutil.copy(os.path.join(GOLDEN, "lens.stinput"), src)
    ctx = hip_lib.st_create_context()
    try:
        assert d.st_amd_read_stinput(ctx, str(src).encode()) == 1
        hip_lib.st_sim_params(ctx, 2000, 100000, 0)
        hip_lib.st_sim_errors(ctx, 1, 1)
        assert stapi.run(hip_lib, ctx, 5) > 0
        assert d.st_amd_write_ray_file(ctx, str(tmp_path / "lib.rays").encode(), 0) == 1
        assert d.st_amd_write_ray_file(ctx, str(tmp_path / "lib.bin").encode(), 1) == 1
        assert d.st_amd_write_ray_file(ctx, str(tmp_path / "bad").encode(), 7) == -1
        res = stapi.fetch_results(hip_lib, ctx)
    finally:
        hip_lib.st_free_context(ctx)
    assert (tmp_path / "lib.rays").read_text() == cli_text
    head, rec = read_binary_rays(tmp_path / "lib.bin")
    assert list(head) == [*res.sun_box, res.nsunrays, len(res)]
    np.testing.assert_array_equal(rec[:, 0:3], res.pos)
    np.testing.assert_array_equal(rec[:, 3:6], res.cos)
    np.testing.assert_array_equal(rec[:, 6], res.stage)
    np.testing.assert_array_equal(rec[:, 7], res.element)
    np.testing.assert_array_equal(rec[:, 8], res.raynum)
