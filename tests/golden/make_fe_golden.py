"""Write tests/golden/fe/: finite-element surface fixtures and the reference's
own strace output on them.

* dish.fed: a structured 31 x 31 mesh of a rippled dish,
  z = (x^2 + y^2) / 8 + 0.003 sin(6x) cos(5y), over [-0.5, 0.5]^2;
* dish_un.fed: 961 nodes of the same surface, 80 on the boundary and the
  rest at seeded random positions;
* dish_dup.fed: dish_un.fed with three nodes repeated (two corners and an
  interior node).  Duplicate nodes make the kriging system singular where
  both are among the nearest seven, and the result there hinges on the last
  bit of pow: a case for the reference build and the oracle, which share
  the host's pow;
* holed.fed: the structured mesh with the nodes inside r < 0.3 removed, so
  rays through the middle reach hash-tree zones that do not exist;
* fe_dish.stinput: both dish facets under a pillbox sun, a receiver
  plane above them (the layout of app/deploy/samples/Surface
  Sample-Finite Element.stinput);
* fe_holed.stinput: the same with holed.fed as the first facet;
* fe_dish.rays: the reference strace (coretrace/strace/main.cpp, built from
  source by `make -C oracle ref` into oracle/_ref/strace_ref) on
  fe_dish.stinput with 300 ray hits, max 200000 sun rays, seed 5, errors on.

Usage: python tests/golden/make_fe_golden.py"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fe")
STRACE_REF = os.path.join(ROOT, "oracle", "_ref", "strace_ref")
ARGS = ["300", "200000", "5", "1", "1", "0"]


def surface(x, y):
    return (x * x + y * y) / 8.0 + 0.003 * np.sin(6.0 * x) * np.cos(5.0 * y)


def write_fed(path, title, x, y):
    z = surface(x, y)
    with open(path, "w") as f:
        f.write(title + "\n%d\n" % len(x))
        for a, b, c in zip(x, y, z):
            f.write("%r %r %r\n" % (float(a), float(b), float(c)))


def structured():
    g = np.linspace(-0.5, 0.5, 31)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    return xx.ravel(), yy.ravel()


def unstructured(n=961):
    rng = np.random.default_rng(20240607)
    edge = np.linspace(-0.5, 0.5, 21)
    inner = edge[1:-1]
    bx = np.concatenate([edge, edge, np.full(19, -0.5), np.full(19, 0.5)])
    by = np.concatenate([np.full(21, -0.5), np.full(21, 0.5), inner, inner])
    m = n - len(bx)
    ix = rng.uniform(-0.5, 0.5, m)
    iy = rng.uniform(-0.5, 0.5, m)
    return np.concatenate([bx, ix]), np.concatenate([by, iy])


HEAD = """# SOLTRACE VERSION 3.2.0 INPUT FILE
SUN\tPTSRC\t0\tSHAPE\tp\tSIGMA\t0\tHALFWIDTH\t4.65
XYZ\t0\t0\t100\tUSELDH\t0\tLDH\t39.5\t90\t10
USER SHAPE DATA\t0
OPTICS LIST COUNT\t2
OPTICAL PAIR\tmirror
OPTICAL.v2\tg\t3\t1\t4\t0.96\t0\t0.95\t1.5\t0.5\t1.2\t1.1\t1.2\t1.3\t1.4\t0\t0\t0\t0
OPTICAL.v2\tg\t3\t1\t4\t0.96\t1\t0.95\t1.5\t0.5\t1.2\t1.1\t1.2\t1.3\t1.4\t0\t0\t0\t0
OPTICAL PAIR\treceiver
OPTICAL.v2\tg\t3\t1\t4\t0\t0\t0.95\t0.2\t1.1\t1.2\t1.1\t1.2\t1.3\t1.4\t0\t0\t0\t0
OPTICAL.v2\tg\t3\t1\t4\t0.96\t1\t0.95\t0.2\t1.1\t1.2\t1.1\t1.2\t1.3\t1.4\t0\t0\t0\t0
STAGE LIST COUNT\t2
STAGE\tXYZ\t0\t0\t0\tAIM\t0\t0\t1\tZROT\t0\tVIRTUAL\t0\tMULTIHIT\t1\tELEMENTS\t2\tTRACETHROUGH\t0
Dishes
"""
FACET = "1\t{x}\t0\t0\t{x}\t0\t1\t0\tr\t1\t1\t0\t0\t0\t0\t0\t0\te\t0\t0\t0\t0\t0\t0\t0\t0\t{fed}\tmirror\t2\t{name}\n"
TAIL = """STAGE\tXYZ\t0\t0\t0\tAIM\t0\t0\t1\tZROT\t0\tVIRTUAL\t0\tMULTIHIT\t1\tELEMENTS\t1\tTRACETHROUGH\t0
Receiver
1\t0.6\t0\t2\t0.6\t0\t-1\t0\tr\t4\t4\t0\t0\t0\t0\t0\t0\tf\t0\t0\t0\t0\t0\t0\t0\t0\t\treceiver\t2
"""


def system(first_fed):
    return (HEAD + FACET.format(x=0, fed=first_fed, name="Structured mesh")
            + FACET.format(x=1.2, fed="dish_un.fed", name="Unstructured mesh") + TAIL)


def main():
    os.makedirs(OUT, exist_ok=True)
    x, y = structured()
    write_fed(os.path.join(OUT, "dish.fed"), "Rippled dish 31 x 31", x, y)
    keep = x * x + y * y >= 0.09
    write_fed(os.path.join(OUT, "holed.fed"), "Rippled dish 31 x 31, hole r < 0.3", x[keep], y[keep])
    ux, uy = unstructured()
    write_fed(os.path.join(OUT, "dish_un.fed"), "Rippled dish, 961 random nodes", ux, uy)
    dup = [0, 20, 500]
    write_fed(os.path.join(OUT, "dish_dup.fed"), "Rippled dish, 961 random nodes and 3 repeated",
              np.concatenate([ux, ux[dup]]), np.concatenate([uy, uy[dup]]))
    with open(os.path.join(OUT, "fe_dish.stinput"), "w") as f:
        f.write(system("dish.fed"))
    with open(os.path.join(OUT, "fe_holed.stinput"), "w") as f:
        f.write(system("holed.fed"))
    if not os.path.exists(STRACE_REF):
        sys.exit("build the reference first: make -C oracle ref")
    with tempfile.TemporaryDirectory() as d:
        for f in ("fe_dish.stinput", "dish.fed", "dish_un.fed"):
            shutil.copy(os.path.join(OUT, f), d)
        subprocess.run([STRACE_REF, "fe_dish.stinput", *ARGS], cwd=d, check=True, capture_output=True)
        shutil.copy(os.path.join(d, "fe_dish.stinput.rays"), os.path.join(OUT, "fe_dish.rays"))
    print("fe_dish.rays", os.path.getsize(os.path.join(OUT, "fe_dish.rays")), "bytes")


if __name__ == "__main__":
    main()
