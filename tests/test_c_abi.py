"""The drop-in boundary, checked without a GPU: the shared library
(coretrace_api.so, no lib prefix, as coretrace/CMakeLists.txt:89-94 names it)
and the static libcoretrace.a export every function include/stapi.h and
include/stapi_amd.h declare; stapi.h declares exactly the reference's 55
entry points (coretrace/stapi.h:68-176); the shared library loads with
ctypes.  No test here calls into the library."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "soltrace_amd", "coretrace_api.so")
AR = os.path.join(ROOT, "soltrace_amd", "libcoretrace.a")
REF_HEADER = "/root/reference/coretrace/stapi.h"


def declared(path):
    """Names of the functions a C header declares (comments stripped)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(st_\w+)\s*\(", text))


def defined_symbols(path, dynamic):
    """Defined function names: C symbols as they are, C++ ones by name."""
    args = ["nm", "--defined-only"] + (["-D"] if dynamic else []) + [path]
    out = subprocess.run(args, capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if re.search(r" T ", line)}
    out = subprocess.run(args[:1] + ["-C"] + args[1:], capture_output=True, text=True, check=True).stdout
    names |= set(re.findall(r" T (st_\w+)\(", out))
    return names


STAPI = declared(os.path.join(ROOT, "include", "stapi.h"))
EXT = declared(os.path.join(ROOT, "include", "stapi_amd.h"))


def test_stapi_header_has_the_55_reference_functions():
    assert len(STAPI) == 55, sorted(STAPI)
    assert not (STAPI & EXT)


@pytest.mark.skipif(not os.path.exists(REF_HEADER), reason="reference sources not present")
def test_stapi_names_equal_the_reference_header():
    ref = declared(REF_HEADER)
    assert STAPI == ref, (sorted(STAPI - ref), sorted(ref - STAPI))


def test_shared_library_exports_every_declaration():
    assert os.path.exists(SO), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    missing = (STAPI | EXT) - defined_symbols(SO, dynamic=True)
    assert not missing, sorted(missing)


def test_static_library_exports_every_declaration():
    assert os.path.exists(AR)
    missing = (STAPI | EXT) - defined_symbols(AR, dynamic=False)
    assert not missing, sorted(missing)


# st_sim_run_data has C++ linkage, as in the reference (coretrace/stapi.cpp:680)
CXX_ONLY = {"st_sim_run_data"}


def test_shared_library_loads_and_resolves():
    lib = ctypes.CDLL(SO)
    for name in sorted((STAPI | EXT) - CXX_ONLY):
        assert getattr(lib, name) is not None


CXX_APP = r"""
#include <vector>
#include "stapi.h"
#include "stapi_amd.h"
int main() {
    st_context_t c = st_create_context();
    std::vector<std::vector<double> > s1, s2;
    int n = st_sim_run_data(c, 1, &s1, &s2, true, 0, 0);
    st_free_context(c);
    return n < 0;
}
"""


def test_cpp_caller_of_st_sim_run_data_links(tmp_path):
    # a C++ caller written against the reference's own prototype links
    # unchanged (linking only: running needs a GPU)
    import shutil
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src, exe = tmp_path / "app.cpp", tmp_path / "app"
    src.write_text(CXX_APP)
    subprocess.run(["g++", "-I", os.path.join(ROOT, "include"), str(src), SO, "-o", str(exe)], check=True)
    assert exe.exists()


APP = r"""
#include <stdio.h>
#include "stapi.h"
#include "stapi_amd.h"
int main(void) {
    st_context_t c = st_create_context();
    if (!c) return 1;
    st_amd_set_precision(c, 32);
    printf("%d\n", st_num_stages(c));
    return st_free_context(c) != 1;
}
"""


@pytest.mark.parametrize("kind", ["static", "shared"])
def test_c_program_links_against_the_library(tmp_path, kind):
    # what the GUI and strace do with the static coretrace library, and any
    # C caller with the shared one (linking only: running needs a GPU)
    import shutil
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("gcc/g++ not available")
    src, obj, exe = tmp_path / "app.c", tmp_path / "app.o", tmp_path / "app"
    src.write_text(APP)
    subprocess.run(["gcc", "-c", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(obj)], check=True)
    if kind == "static":
        cmd = ["g++", str(obj), AR, "-L/opt/rocm/lib", "-lamdhip64", "-o", str(exe)]
    else:
        cmd = ["gcc", str(obj), SO, "-o", str(exe)]
    subprocess.run(cmd, check=True)
    assert exe.exists()
