"""The C-ABI library loads and exports every symbol include/*.h declares (CPU only: no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

import cpmgpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(cpm_[A-Za-z0-9_]+)\s*\(", txt):
            syms.add(m.group(1))
    return syms


def test_library_exports_declared_symbols():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = ctypes.CDLL(cpmgpu.LIB_PATH)
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(cpmgpu.EXPORTS) <= syms


def test_default_opts_match_reference():
    o = cpmgpu.opts()
    # src/Solver.cpp:110: tolerance 1e-10, at most 100000 iterations
    assert o.tol == 1e-10 and o.max_iters == 100000 and o.precision == cpmgpu.CPM_FP64


def test_cpp_facade_compiles(tmp_path):
    """The C++ facade (cpm::Solver and the problem classes) and the drop-in test programs compile
    against the C ABI header."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    for prog in ("solver_dropin", "problems_dropin"):
        src = os.path.join(REPO, "tests", "cpp", prog + ".cpp")
        out = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I",
                              os.path.join(REPO, "include"), "-I", os.path.join(REPO, "cpm-ibc_amd", "host"), src],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, (prog, out.stderr)
