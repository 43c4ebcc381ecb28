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

from test_strace import SAMPLES, STRACE, STRACE_ORACLE, run_cli

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
