"""Write tests/golden/stinput/*.rays: the reference's own strace
(coretrace/strace/main.cpp, built from source by `make -C oracle ref` into
oracle/_ref/strace_ref) on the sample systems copied here from
app/deploy/samples, 100 ray hits, max 100000 sun rays, seed 7, sunshape and
optics errors on, power-tower grouping off.

Usage: python tests/golden/make_stinput_golden.py"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
STRACE_REF = os.path.join(ROOT, "oracle", "_ref", "strace_ref")
SAMPLES = ["lens", "transmissivity-table", "Refraction", "Power-tower-surround_singlefacet_large_afternoon"]
ARGS = ["100", "100000", "7", "1", "1", "0"]


def main():
    if not os.path.exists(STRACE_REF):
        sys.exit("build the reference first: make -C oracle ref")
    out_dir = os.path.join(HERE, "stinput")
    with tempfile.TemporaryDirectory() as d:
        for name in SAMPLES:
            src = os.path.join(out_dir, name + ".stinput")
            tmp = os.path.join(d, name + ".stinput")
            shutil.copy(src, tmp)
            subprocess.run([STRACE_REF, tmp, *ARGS], cwd=d, check=True, capture_output=True)
            shutil.copy(tmp + ".rays", os.path.join(out_dir, name + ".rays"))
            print(name, os.path.getsize(os.path.join(out_dir, name + ".rays")), "bytes")


if __name__ == "__main__":
    main()
