"""C-ABI boundary checks that need no GPU: the library loads, exports every symbol the
public header declares, and its host-side argument handling follows the reference driver
(mgemm_parallel.cpp:66 early return) or fails loudly where the reference would not check."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import mwblas

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names += re.findall(r"MWGEMM_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_dropin_symbols():
    syms = header_symbols()
    for s in ("mgemm_", "mgemm_64_", "wgemm_", "wgemm_64_", "mgemm_serial", "wgemm_serial",
              "mgemmtr_", "mgemmtr_64_", "wgemmtr_", "wgemmtr_64_", "mgemmtr_serial",
              "wgemmtr_serial", "mgemv_", "mgemv_64_", "wgemv_", "wgemv_64_"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = mwblas.lib()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    out = os.popen(f"nm -D --defined-only {mwblas.LIB_PATH}").read()
    exported = set(re.findall(r" T (\S+)", out))
    assert set(header_symbols()) <= exported
    # internal C++ symbols stay hidden
    assert not any(e.startswith("_ZN2mw") for e in exported)


def test_empty_problem_returns_without_touching_anything():
    C = np.full((3, 4, 2), 7.0)
    mwblas.mgemm("N", "N", 0, 3, 5, [1, 0], None, 1, None, 1, [0, 0], C, 4)
    mwblas.mgemm("N", "N", 4, 0, 5, [1, 0], None, 4, None, 5, [0, 0], C, 4)
    mwblas.wgemm("N", "N", 0, 0, 5, [1, 0, 0, 0], None, 1, None, 1, [0, 0, 0, 0], C, 4)
    assert np.all(C == 7.0)


def test_invalid_arguments_fail_loudly():
    A = np.zeros((4, 4, 2))
    C = np.zeros((4, 4, 2))
    # drop-in entry points: a leading dimension shorter than its rows is rejected before
    # anything is staged (the reference would read overlapping columns); letters are read
    # as the reference reads them (tests/test_gpu_letters.py), never rejected
    with pytest.raises(mwblas.MwgemmError, match="leading dimension"):
        mwblas.mgemm("N", "N", 4, 4, 4, [1, 0], A, 3, A, 4, [0, 0], C, 4)
    with pytest.raises(mwblas.MwgemmError, match="leading dimension"):
        mwblas.mgemm("X", "N", 4, 4, 4, [1, 0], A, 3, A, 4, [0, 0], C, 4)   # 'X' reads as 'T'
    with pytest.raises(mwblas.MwgemmError, match="leading dimension"):
        mwblas.wgemmtr("L", "N", "N", 4, 4, [1, 0, 0, 0], A, 4, A, 4, [0, 0, 0, 0], C, 3)
    mwblas.mgemmtr("U", "N", "N", 0, 4, [1, 0], A, 4, A, 4, [0, 0], C, 4)   # n == 0: no-op
    v = np.zeros((8, 2))
    with pytest.raises(mwblas.MwgemmError, match="leading dimension"):
        mwblas.mgemv("N", 4, 4, [1, 0], A, 3, v, 1, [0, 0], v, 1)
    with pytest.raises(mwblas.MwgemmError, match="increment"):
        mwblas.wgemv("T", 4, 4, [1, 0, 0, 0], A, 4, v, 0, [0, 0, 0, 0], v, 1)
    y = np.full((4, 2), 3.0)
    mwblas.mgemv("N", 0, 4, [1, 0], A, 1, v, 1, [2, 0], y, 1)     # m == 0: y untouched
    assert np.all(y == 3.0)
    assert np.all(C == 0.0)
    # the library's own device entry points (no reference counterpart) take only N/T/C and
    # U/L; the letter is checked before any pointer is used
    with pytest.raises(mwblas.MwgemmError, match="trans"):
        mwblas.gemm_device(False, "X", "N", 4, 4, 4, [1, 0], None, 4, None, 4, [0, 0], None, 4)
    with pytest.raises(mwblas.MwgemmError, match="trans"):
        mwblas.gemm_device(True, "N", "Q", 4, 4, 4, [1, 0, 0, 0], None, 4, None, 4,
                           [0, 0, 0, 0], None, 4)
    with pytest.raises(mwblas.MwgemmError, match="uplo"):
        mwblas.gemmtr_device(False, "X", "N", "N", 4, 4, [1, 0], None, 4, None, 4, [0, 0],
                             None, 4)
    with pytest.raises(mwblas.MwgemmError, match="trans"):
        mwblas.gemmtr_device(True, "U", "N", "Z", 4, 4, [1, 0, 0, 0], None, 4, None, 4,
                             [0, 0, 0, 0], None, 4)
    with pytest.raises(mwblas.MwgemmError, match="trans"):
        mwblas.gemv_device(False, "Q", 4, 4, [1, 0], None, 4, None, 1, [0, 0], None, 1)


def test_path_switch():
    prev = mwblas.set_path(mwblas.PATH_OZAKI)
    assert mwblas.lib().mwgemm_get_path() == mwblas.PATH_OZAKI
    mwblas.set_path(prev)
    with pytest.raises(ValueError):
        mwblas.set_path(7)


def test_abort_mode_is_the_default_for_native_callers():
    """A C caller that never touches the error API gets an abort on failure; run it in a
    child process (no GPU needed: the failure is an argument error)."""
    code = (
        "import ctypes, sys; L = ctypes.CDLL(sys.argv[1]);"
        "i = ctypes.c_int(4); bad = ctypes.c_int(1); d = (ctypes.c_double * 64)(); d[0] = 1.0;"
        "L.mgemm_(b'N', b'N', ctypes.byref(i), ctypes.byref(i), ctypes.byref(i), d, d,"
        " ctypes.byref(bad), d, ctypes.byref(i), d, d, ctypes.byref(i))"
    )
    import subprocess
    import sys
    env = dict(os.environ)
    env.pop("MWGEMM_ON_ERROR", None)
    r = subprocess.run([sys.executable, "-c", code, mwblas.LIB_PATH], capture_output=True,
                       text=True, env=env)
    assert r.returncode != 0
    assert "leading dimension" in r.stderr
    env["MWGEMM_ON_ERROR"] = "return"
    r = subprocess.run([sys.executable, "-c", code, mwblas.LIB_PATH], capture_output=True,
                       text=True, env=env)
    assert r.returncode == 0 and "leading dimension" in r.stderr
