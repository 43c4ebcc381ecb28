"""Pins the oracle's random streams (CPU only).

* RefStream, the oracle's restatement of the reference's sequential stream
  (coretrace/mtrand.h:298-348: std::default_random_engine feeding
  uniform_real_distribution<double>(0,1) and normal_distribution<double>(0,1)),
  is compared draw for draw with the std:: types themselves, built by g++
  from a probe compiled inside the test.
* Philox4x32-10 is checked against the Random123 known-answer vectors
  (Salmon et al., SC'11; kat_vectors of the Random123 distribution).
* The counter stream's layout: uniform k of stream (seed, stage, id) is
  word k%4 of the Philox block with counter (id_lo, id_hi, stage, k/4) and
  key (seed, 0x53544F52), times 2^-32.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "st_oracle.so")

PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char **argv) {
    const unsigned seed = (unsigned)strtoul(argv[1], 0, 10);
    const int n = atoi(argv[2]), normals = atoi(argv[3]);
    std::default_random_engine e(seed);
    std::uniform_real_distribution<double> u(0., 1.);
    std::normal_distribution<double> g(0., 1.);
    for (int i = 0; i < n; ++i) printf("%a\n", normals ? 0.0 + g(e) * 1.0 : u(e));
    return 0;
}
"""

SEEDS = [0, 1, 10, 123, 2147483646, 2147483647, 4294967295]


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(ORACLE_SO):
        pytest.skip("oracle not built (make -C oracle)")
    d = C.CDLL(ORACLE_SO)
    d.st_oracle_ref_uniforms.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_double)]
    d.st_oracle_ref_normals.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_double)]
    d.st_oracle_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    d.st_oracle_counter_uniform.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64]
    d.st_oracle_counter_uniform.restype = C.c_double
    return d


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("rngprobe")
    src, exe = d / "probe.cpp", d / "probe"
    src.write_text(PROBE)
    subprocess.run([gxx, "-O2", "-std=c++11", str(src), "-o", str(exe)], check=True)
    return str(exe)


def _std_draws(probe, seed, n, normals):
    out = subprocess.run([probe, str(seed), str(n), str(int(normals))], capture_output=True, text=True,
                         check=True).stdout.split()
    return np.array([float.fromhex(v) for v in out])


@pytest.mark.parametrize("seed", SEEDS)
def test_ref_uniforms_equal_libstdcxx(lib, probe, seed):
    n = 5000
    buf = (C.c_double * n)()
    lib.st_oracle_ref_uniforms(seed, n, buf)
    ours = np.frombuffer(buf, dtype=np.float64)
    np.testing.assert_array_equal(ours, _std_draws(probe, seed, n, False))


@pytest.mark.parametrize("seed", SEEDS)
def test_ref_normals_equal_libstdcxx(lib, probe, seed):
    n = 5000
    buf = (C.c_double * n)()
    lib.st_oracle_ref_normals(seed, n, buf)
    ours = np.frombuffer(buf, dtype=np.float64)
    np.testing.assert_array_equal(ours, _std_draws(probe, seed, n, True))


# Random123 known-answer vectors for philox4x32_10: counter, key, output
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(lib, ctr, key, want):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.st_oracle_philox(c, k, o)
    assert tuple(o) == want


def test_counter_stream_layout(lib):
    rng = np.random.default_rng(5)
    for _ in range(200):
        seed = int(rng.integers(0, 2 ** 32))
        stage = int(rng.integers(0, 16))
        sid = int(rng.integers(0, 2 ** 40))
        k = int(rng.integers(0, 5000))
        c = (C.c_uint32 * 4)(sid & 0xFFFFFFFF, sid >> 32, stage, k // 4)
        key = (C.c_uint32 * 2)(seed, 0x53544F52)
        o = (C.c_uint32 * 4)()
        lib.st_oracle_philox(c, key, o)
        assert lib.st_oracle_counter_uniform(seed, stage, sid, k) == o[k % 4] * 2.0 ** -32


def test_counter_uniforms_are_uniform(lib):
    # 40,000 draws of one stream: mean, variance and a 20-bin chi^2
    u = np.array([lib.st_oracle_counter_uniform(7, 0, 12345, k) for k in range(40000)])
    assert abs(u.mean() - 0.5) < 4 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.003
    counts, _ = np.histogram(u, bins=20, range=(0, 1))
    chi2 = ((counts - u.size / 20) ** 2 / (u.size / 20)).sum()
    assert chi2 < 45.3  # p = 0.001 for 19 degrees of freedom
