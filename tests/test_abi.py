"""The product library loads and exports exactly the C ABI the headers
declare.  CPU only: no call here reaches a device kernel."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

import l3cases as L

HEADERS = ["cblas.h", "mublis.h", "mublis_instantiate.h", "mublis_gfx950.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(L.ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b((?:cblas|mublis)_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    import mublis_amd
    if not os.path.exists(mublis_amd.LIB_PATH):
        mublis_amd.build()
    return mublis_amd.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    # the 12 CBLAS routines + xerbla, 12 mublis_* + 2 drivers, the config hook
    assert len([n for n in names if n.startswith("cblas_")]) == 13
    assert "mublis_get_context" in names and "mublis_l3_ddriver" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_xerbla_is_weak():
    out = subprocess.run(["nm", "-D", os.path.join(L.ROOT, "mublis_amd", "lib",
                                                    "libmublis.so")],
                         capture_output=True, text=True).stdout
    line = [ln for ln in out.splitlines() if ln.endswith(" cblas_xerbla")][0]
    assert line.split()[1] == "W"


def test_enum_values_match_reference():
    """include/cblas.h:18-22 and include/mublis.h:8-45 of the reference."""
    text = open(os.path.join(L.ROOT, "include", "cblas.h")).read()
    for name, val in [("CblasRowMajor", 101), ("CblasColMajor", 102), ("CblasNoTrans", 111),
                      ("CblasTrans", 112), ("CblasUpper", 121), ("CblasLower", 122),
                      ("CblasNonUnit", 131), ("CblasUnit", 132), ("CblasLeft", 141),
                      ("CblasRight", 142)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", text), name
    text = open(os.path.join(L.ROOT, "include", "mublis.h")).read()
    for name, val in [("MUBLIS_TRANSPOSE", 1), ("MUBLIS_LOWER", 1), ("MUBLIS_UPPER", 2),
                      ("MUBLIS_RIGHT", 1), ("MUBLIS_UNIT_DIAG", 1),
                      ("MUBLIS_STRUC_SYMMETRIC", 1), ("MUBLIS_STRUC_TRIANGULAR", 2),
                      ("MUBLIS_PACKM_DIAG_INVERT", 2), ("MUBLIS_L3_UPPER", 2)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", text), name


def test_empty_problems_return_without_a_device(lib):
    """m == 0 or n == 0 returns 0 before any device work
    (frame/l3/l3_driver.c:61-62, frame/l3/trsm.c:24-25)."""
    assert lib.mublis_dgemm(0, 0, 0, 5, 5, 1.0, None, 1, 1, None, 1, 1, 0.0, None, 1, 1) == 0
    assert lib.mublis_strsm(0, 1, 0, 0, 5, 0, 1.0, None, 1, 1, None, 1, 1) == 0


def test_replay_rejects_bad_arguments_without_a_device(lib):
    """mublis_server_replay checks its arguments before touching a server:
    no server, a negative count, or missing arrays return MUBLIS_EINVAL and
    report nothing submitted."""
    from mublis_amd import EINVAL
    done, lag = ctypes.c_int(7), ctypes.c_double(3.0)
    assert lib.mublis_server_replay(None, None, None, 0, None, ctypes.byref(done),
                                    ctypes.byref(lag)) == EINVAL
    assert done.value == 0 and lag.value == 0.0
    fake = ctypes.c_void_p(1)  # never dereferenced: the count is checked first
    assert lib.mublis_server_replay(fake, None, None, -1, None, None, None) == EINVAL
    assert lib.mublis_server_replay(fake, None, None, 2, None, None, None) == EINVAL
    # offsets outside [0, 86400] s are refused before the server is touched
    for bad in (-1.0, float("nan"), float("inf"), 1e9):
        at = (ctypes.c_double * 1)(bad)
        tk = (ctypes.c_uint64 * 1)()
        req = ctypes.create_string_buffer(256)
        assert lib.mublis_server_replay(fake, req, at, 1, tk, None, None) == EINVAL, bad


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="needs a host without a GPU")
def test_no_device_reports_backend_failure_through_xerbla():
    """Without a gfx950 device every call fails with status 1 and CBLAS
    reports it through cblas_xerbla(0, ...), which exits
    (frame/compat/cblas/cblas.c:100-112)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np, mublis_amd as M\n"
        "lib = M.load()\n"
        "a = np.ones((4, 4)); c = np.zeros((4, 4))\n"
        "print('status', lib.mublis_dgemm(0, 0, 4, 4, 4, 1.0, M.ptr(a), 1, 4, M.ptr(a), 1, 4,"
        " 0.0, M.ptr(c), 1, 4), flush=True)\n"
        "lib.cblas_dgemm(102, 111, 111, 4, 4, 4, 1.0, M.ptr(a), 4, M.ptr(a), 4, 0.0,"
        " M.ptr(c), 4)\n"
        "print('RETURNED')\n" % L.ROOT)
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=120)
    assert "status 1" in p.stdout
    assert "RETURNED" not in p.stdout and p.returncode == 1
    assert "Routine cblas_dgemm failed" in p.stderr
    assert "MuBLIS backend returned status 1" in p.stderr
