"""The .stinput reader and .rays writer (soltrace_amd/strace), checked on the
CPU.  The reader and writer are the product's own source; run on the oracle
engine in the reference's stream mode (oracle/_build/strace_oracle), their
.rays file must equal, byte for byte, what the reference's own strace
(coretrace/strace/main.cpp) writes for the same sample system and arguments:
against the reference build itself when it is present, and always against
the golden files it wrote (tests/golden/make_stinput_golden.py).  The samples
are app/deploy/samples systems: a two-stage lens, a transmissivity table, a
refraction case with a user sun table, and a 2,400-heliostat SolarPILOT field
with a Buie-like sun and a cylindrical receiver."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "stinput")
STRACE_ORACLE = os.path.join(ROOT, "oracle", "_build", "strace_oracle")
STRACE_REF = os.path.join(ROOT, "oracle", "_ref", "strace_ref")
STRACE = os.path.join(ROOT, "soltrace_amd", "strace", "strace")
SAMPLES = ["lens", "transmissivity-table", "Refraction", "Power-tower-surround_singlefacet_large_afternoon"]
GOLDEN_ARGS = ["100", "100000", "7", "1", "1", "0"]

needs_oracle = pytest.mark.skipif(not os.path.exists(STRACE_ORACLE), reason="oracle not built (make -C oracle)")


FE_GOLDEN = os.path.join(ROOT, "tests", "golden", "fe")
FE_FILES = ("dish.fed", "dish_un.fed", "dish_dup.fed", "holed.fed")


def run_cli(exe, tmp_path, name, args, prefix=(), src_dir=GOLDEN, extra=()):
    """Run a strace command line on a fresh copy of a sample (and of the
    surface files it names); return the .rays text (None when none was
    written) and the process result."""
    src = os.path.join(src_dir, name + ".stinput")
    dst = tmp_path / (name + ".stinput")
    shutil.copy(src, dst)
    for f in extra:
        shutil.copy(os.path.join(src_dir, f), tmp_path / f)
    r = subprocess.run([exe, *prefix, str(dst), *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    rays = tmp_path / (name + ".stinput.rays")
    return (rays.read_text() if rays.exists() else None), r


@needs_oracle
@pytest.mark.parametrize("name", SAMPLES)
def test_rays_equal_reference_golden(tmp_path, name):
    got, r = run_cli(STRACE_ORACLE, tmp_path, name, GOLDEN_ARGS)
    assert r.returncode == 0, r.stdout + r.stderr
    assert got == open(os.path.join(GOLDEN, name + ".rays")).read()


@needs_oracle
@pytest.mark.skipif(not os.path.exists(STRACE_REF), reason="reference build absent (make -C oracle ref)")
@pytest.mark.parametrize("name", SAMPLES)
@pytest.mark.parametrize("args", [["1500", "300000", "12", "1", "1", "0"], ["800", "300000", "3", "0", "0", "1"]],
                         ids=["errors", "no_errors_power_tower"])
def test_rays_equal_reference_build(tmp_path, name, args):
    a, b = tmp_path / "ours", tmp_path / "ref"
    a.mkdir()
    b.mkdir()
    ours, r1 = run_cli(STRACE_ORACLE, a, name, args)
    ref, r2 = run_cli(STRACE_REF, b, name, args)
    assert r1.returncode == 0 and r2.returncode == 0, r1.stdout + r2.stdout
    assert ours is not None and ours == ref


def _header(text):
    return [float(v) for v in text.splitlines()[1].split()]


@needs_oracle
def test_rays_header_counts(tmp_path):
    text, _ = run_cli(STRACE_ORACLE, tmp_path, "lens", GOLDEN_ARGS)
    lines = text.splitlines()
    head = _header(text)
    assert len(lines) - 3 == int(head[5])                 # Length rows follow the two comment lines
    assert head[0] < head[1] and head[2] < head[3]        # sun box
    rows = [ln.split() for ln in lines[3:]]
    assert all(len(x) == 9 for x in rows)
    assert max(int(x[8]) for x in rows) <= int(head[4])  # ray numbers within the sun rays drawn


def test_cli_usage_and_errors(tmp_path):
    if not os.path.exists(STRACE):
        pytest.skip("build first")
    r = subprocess.run([STRACE], capture_output=True, text=True)
    assert r.returncode == 255 and "usage" in r.stdout
    r = subprocess.run([STRACE, str(tmp_path / "missing.stinput"), "10", "100", "1", "1", "1"], capture_output=True,
                       text=True)
    assert r.returncode == 255 and "failed to open" in r.stdout


@needs_oracle
def test_reader_rejects_headerless_and_short_lines(tmp_path):
    # the strace reader requires the '#' version line and 29 element fields
    # (coretrace/strace/main.cpp:258-264, 364-375)
    src = open(os.path.join(GOLDEN, "lens.stinput")).read()
    bad = tmp_path / "noheader.stinput"
    bad.write_text(src.split("\n", 1)[1])
    r = subprocess.run([STRACE_ORACLE, str(bad), "10", "1000", "1", "1", "1", "0"], capture_output=True, text=True)
    assert r.returncode == 255 and "must start with '#'" in r.stdout
    lines = src.split("\n")
    k = next(i for i, ln in enumerate(lines) if ln.startswith("1\t0\t0\t0.5"))
    lines[k] = "\t".join(lines[k].split("\t")[:20])
    short = tmp_path / "short.stinput"
    short.write_text("\n".join(lines))
    r = subprocess.run([STRACE_ORACLE, str(short), "10", "1000", "1", "1", "1", "0"], capture_output=True, text=True)
    assert r.returncode == 255 and "too few tokens for element" in r.stdout


def read_binary_rays(path):
    """The GUI's binary ray file (app/src/project.cpp:915-988): six doubles
    of header, then nine doubles per record."""
    import numpy as np
    raw = np.fromfile(path, dtype="<f8")
    head = raw[:6]
    n = int(head[5])
    assert raw.size == 6 + 9 * n
    return head, raw[6:].reshape(n, 9)


@needs_oracle
def test_binary_ray_file_matches_text(tmp_path):
    import numpy as np
    shutil.copy(os.path.join(GOLDEN, "Refraction.stinput"), tmp_path / "r.stinput")
    r = subprocess.run([STRACE_ORACLE, "--binary", str(tmp_path / "r.stinput"), *GOLDEN_ARGS], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    head, rec = read_binary_rays(tmp_path / "r.stinput.bin")
    text = (tmp_path / "r.stinput.rays").read_text().splitlines()
    thead = np.array([float(v) for v in text[1].split()])
    trow = np.array([[float(v) for v in ln.split()] for ln in text[3:]])
    # the text carries 7 significant digits of the same doubles
    np.testing.assert_allclose(head, thead, rtol=6e-7, atol=0)
    assert rec.shape == trow.shape and head[5] == len(trow) > 0
    np.testing.assert_array_equal(rec[:, 6:], trow[:, 6:])
    np.testing.assert_allclose(rec[:, :6], trow[:, :6], rtol=6e-7, atol=1e-300)


REF_SAMPLES = "/root/reference/app/deploy/samples"


@needs_oracle
@pytest.mark.skipif(not (os.path.exists(STRACE_REF) and os.path.isdir(REF_SAMPLES)),
                    reason="reference build and samples absent")
@pytest.mark.parametrize("args", [["600", "200000", "4", "1", "1", "0"], ["300", "200000", "9", "0", "0", "0"]])
def test_finite_element_sample_equals_reference_build(tmp_path, args):
    # the reference's own finite-element sample: two facets meshed by 10,000
    # nodes each (a structured and an unstructured mesh), kriging over the
    # hash-tree neighbourhood (coretrace/vshot.cpp:166-238, treemesh.cpp)
    name = "Surface Sample-Finite Element"
    outs = []
    for exe, sub in ((STRACE_ORACLE, "ours"), (STRACE_REF, "ref")):
        d = tmp_path / sub
        d.mkdir()
        for f in (name + ".stinput", "sample.fed", "sample_un.fed"):
            shutil.copy(os.path.join(REF_SAMPLES, f), d / f)
        r = subprocess.run([exe, name + ".stinput", *args], cwd=d, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:]
        outs.append((d / (name + ".stinput.rays")).read_text())
    assert outs[0] == outs[1] and len(outs[0].splitlines()) > 300


@needs_oracle
def test_finite_element_fixture_equals_reference_golden(tmp_path):
    # a rippled dish meshed twice (structured; random with duplicate corner
    # nodes), golden .rays written by the reference build
    # (tests/golden/make_fe_golden.py)
    got, r = run_cli(STRACE_ORACLE, tmp_path, "fe_dish", ["300", "200000", "5", "1", "1", "0"], src_dir=FE_GOLDEN,
                     extra=FE_FILES)
    assert r.returncode == 0, r.stdout
    assert got == open(os.path.join(FE_GOLDEN, "fe_dish.rays")).read()


@needs_oracle
@pytest.mark.skipif(not os.path.exists(STRACE_REF), reason="reference build absent (make -C oracle ref)")
@pytest.mark.parametrize("name", ["fe_dish", "fe_holed"])
def test_finite_element_fixture_equals_reference_build(tmp_path, name):
    args = ["700", "400000", "11", "1", "1", "0"]
    (tmp_path / "ours").mkdir()
    (tmp_path / "ref").mkdir()
    ours, r1 = run_cli(STRACE_ORACLE, tmp_path / "ours", name, args, src_dir=FE_GOLDEN, extra=FE_FILES)
    ref, r2 = run_cli(STRACE_REF, tmp_path / "ref", name, args, src_dir=FE_GOLDEN, extra=FE_FILES)
    assert r1.returncode == r2.returncode == 0
    assert ours == ref
    assert r1.stdout.splitlines()[-1] == r2.stdout.splitlines()[-1]


@needs_oracle
def test_finite_element_hole_fails_the_run(tmp_path):
    # rays through the mesh's hole reach hash-tree zones that do not exist;
    # the reference throws out of Trace (coretrace/vshot.cpp:177-178) and
    # strace reports it and writes no .rays file
    got, r = run_cli(STRACE_ORACLE, tmp_path, "fe_holed", ["300", "200000", "5", "1", "1", "0"], src_dir=FE_GOLDEN,
                     extra=FE_FILES)
    assert got is None
    assert r.stdout.splitlines()[-1] == "error: trace error: Interpolation failed to find nearby nodes to ray hit " \
                                        "in FEA surface."
