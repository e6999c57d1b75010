"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and reports errors as status codes (no GPU work)."""

import ctypes
import os
import re

import numpy as np
import pytest

from fastoverlap_amd import _lib
from fastoverlap_amd.batch import compact_index, compact_to_dense, dense_to_compact, kcount

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    with open(os.path.join(ROOT, "include", "fastoverlap_hip.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(fo_\w+)\(", txt, re.M)))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding declares each one
    assert set(syms) <= set(_lib.exported_symbols())


def test_status_codes_without_gpu_work():
    lib = _lib.load()
    assert lib.fo_version() >= 10000
    # NULL arguments are rejected with FO_ERR_INVALID, never a crash
    assert lib.fo_context_create(0, 0, None) == 1
    assert lib.fo_sph_batch_run(None, 31, 0) == 1
    assert lib.fo_per_batch_sync(None) == 1
    assert b"NULL" in lib.fo_last_error()
    assert lib.fo_sph_batch_destroy(None) == 0


def test_fortran_shim_compiles():
    import shutil
    import subprocess
    import tempfile
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        pytest.skip("amdflang not available")
    src = os.path.join(ROOT, "fastoverlap_amd", "fortran", "fastoverlap_hip.f90")
    if not os.path.exists(src):
        pytest.skip("Fortran shim source not present")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([flang, "-c", src, "-o", os.path.join(d, "f.o"), "-J", d],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_compact_layout_roundtrip():
    L = 5
    l, m1, m2 = compact_index(L)
    assert len(l) == kcount(L)
    # kidx(l,m1,m2) = l(4l^2-1)/3 + (m1+l)(2l+1) + (m2+l)
    k = l * (4 * l * l - 1) // 3 + (m1 + l) * (2 * l + 1) + (m2 + l)
    np.testing.assert_array_equal(k, np.arange(kcount(L)))
    c = np.arange(kcount(L)) * (1 + 2j)
    np.testing.assert_array_equal(dense_to_compact(compact_to_dense(c, L)), c)
