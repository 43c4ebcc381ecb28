"""The tabulated ('d') sunshape's polar-angle law in the two random streams
of the oracle.  The reference's stream (mode 0) runs the reference's
rejection loop (coretrace/errors.cpp:160-185; the oracle is bit-exact with
the reference build there); the per-ray counter streams the GPU uses (mode
1) draw the same law by inverting its CDF, one 53-bit uniform per ray.
Both are held to the law itself -- density proportional to I(theta) theta
on [0, max angle], I the table's linear interpolation, integrated here
numerically on a fine grid -- by one-sample Kolmogorov-Smirnov tests, and
to each other by a two-sample test, on the Buie CSR 0.1 table and on a
coarse made-up table whose segments the inversion must follow exactly."""
import math

import numpy as np
import pytest
from scipy import stats

import physics_checks as P
from harness import counter_mode
from soltrace_amd import workloads

TABLES = {
    "buie_csr0.1": workloads.buie_profile(0.1),
    # few, uneven segments, a zero-intensity stretch and a rise
    "coarse": (np.array([0.0, 1.5, 4.0, 4.5, 7.0, 9.0]), np.array([1.0, 0.9, 0.2, 0.0, 0.0, 0.3])),
}


def law_cdf(ang, inten):
    grid = np.linspace(0.0, ang[-1], 400001)
    dens = np.interp(grid, ang, inten) * grid
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    return grid, cdf / cdf[-1]


def _theta(lib, table, mode):
    setup = counter_mode if mode == 1 else None
    return P._sun_spread(lib, setup, "d", 0.0, table=table, n=20000) * 1000.0  # mrad


@pytest.mark.parametrize("name", sorted(TABLES))
def test_both_streams_follow_the_law(oracle_lib, name):
    ang, inten = TABLES[name]
    grid, cdf = law_cdf(ang, inten)
    samples = {m: _theta(oracle_lib, (ang, inten), m) for m in (0, 1)}
    for m, th in samples.items():
        assert th.max() <= ang[-1] * (1 + 1e-9)
        ks = stats.kstest(th, lambda x: np.interp(x, grid, cdf))
        assert ks.pvalue > 1e-3, (m, ks)
    ks2 = stats.ks_2samp(samples[0], samples[1])
    assert ks2.pvalue > 1e-3, ks2


def test_a_wrong_law_is_rejected(oracle_lib):
    # the same test on the counter stream's samples refuses the law without
    # the polar area factor theta (a plausible mistake: density I(theta))
    ang, inten = TABLES["buie_csr0.1"]
    th = _theta(oracle_lib, (ang, inten), 1)
    grid = np.linspace(0.0, ang[-1], 400001)
    dens = np.interp(grid, ang, inten)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    assert stats.kstest(th, lambda x: np.interp(x, grid, cdf)).pvalue < 1e-6
