"""GPU parity of the generic driver with caller-built descriptors
(mublis_l3_?driver, include/mublis.h:41-83), called through the C ABI.

The descriptors (tests/l3cases.py DRIVER_CASES) are the ones the oracle is
pinned on against both reference builds' own mublis_l3_?driver: trapezoidal
TRIANGULAR operands (m != k) with NONUNIT, UNIT and INVERT diagonals, offset
SYMMETRIC operands on either side, ji masks on rectangular C, jp / pi equal
to the triangular operand's uplo, and strides with no unit stride.  The two
DRIVER_MASK_ONLY descriptors put jp / pi on GENERAL operands, where the
per-term definition holds (the reference's own result there depends on its
cache-block sizes).

Bar: the written elements within 2 (k + 2) eps (|alpha| ||Av|| ||Bv|| +
|beta| ||C0||) of the float64 value and within twice that of the oracle's
result; everything else in C's storage (the ji-excluded part, column gaps)
bit-identical to the input; NaN wherever a view must not read (the unstored
triangle, a unit diagonal, elements outside jp / pi, C when beta == 0).
"""
import ctypes

import numpy as np
import pytest

import l3cases as L
import mublis_amd as M

pytestmark = pytest.mark.gpu

CASES = L.DRIVER_CASES + L.DRIVER_MASK_ONLY


@pytest.fixture(scope="module")
def lib():
    lib = M.load()
    assert lib.mublis_rt_init() == 0, lib.mublis_rt_last_error()
    return lib


@pytest.fixture(scope="module")
def orc_lib():
    return ctypes.CDLL(L.ORACLE_LIB)


def _oracle(orc_lib, p, dt):
    ct = ctypes.c_float if np.dtype(dt) == np.float32 else ctypes.c_double
    f = getattr(orc_lib, "orc_sdriver" if ct is ctypes.c_float else "orc_ddriver")
    f.argtypes = [L._I] * 3 + [ct] + [L._VP, L._I, L._I] * 2 + [ct, L._VP, L._I, L._I] + \
        [L._I] * 9
    f.restype = L._I
    out = p["c"].copy()
    assert f(p["m"], p["n"], p["k"], p["alpha"], p["a"].ctypes.data, p["rs_a"], p["cs_a"],
             p["b"].ctypes.data, p["rs_b"], p["cs_b"], p["beta"], out.ctypes.data, p["rs_c"],
             p["cs_c"], *p["desc"]) == 0
    return out


def _check(p, got, want, dt, who):
    idx, w = p["idx"], p["write"]
    res = got[idx].astype(np.float64)
    bound = L.driver_bound(p, dt)
    err = np.linalg.norm(np.where(w, res - p["exact"], 0))
    assert np.isfinite(err) and err <= bound, (who, err, bound)
    diff = np.linalg.norm(np.where(w, res - want[idx].astype(np.float64), 0))
    assert diff <= 2 * bound, (who, diff, bound)
    keep = np.ones(got.size, bool)
    keep[idx[w]] = False
    np.testing.assert_array_equal(got[keep], p["c"][keep], err_msg=who)


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_driver_host_pointers(lib, orc_lib, dt, case):
    """Host operands: staged over their whole strided footprint (for an offset
    SYMMETRIC view the box its reflected reads cover) and copied back."""
    p = L.driver_problem(case, dt)
    want = _oracle(orc_lib, p, dt)
    f = L.driver_fn(lib, dt)
    got = p["c"].copy()
    desc = L.Product(*p["desc"])
    st = f(p["m"], p["n"], p["k"], p["alpha"], p["a"].ctypes.data, p["rs_a"], p["cs_a"],
           p["b"].ctypes.data, p["rs_b"], p["cs_b"], p["beta"], got.ctypes.data, p["rs_c"],
           p["cs_c"], ctypes.byref(desc))
    assert st == 0
    _check(p, got, want, dt, case[0])


@pytest.mark.parametrize("dt", L.DTYPES)
@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_driver_device_pointers(lib, orc_lib, dt, case):
    """Device operands are used in place (no staging)."""
    import torch
    p = L.driver_problem(case, dt)
    want = _oracle(orc_lib, p, dt)
    f = L.driver_fn(lib, dt)
    da = torch.from_numpy(p["a"]).cuda()
    db = torch.from_numpy(p["b"]).cuda()
    dc = torch.from_numpy(p["c"].copy()).cuda()
    desc = L.Product(*p["desc"])
    st = f(p["m"], p["n"], p["k"], p["alpha"], da.data_ptr(), p["rs_a"], p["cs_a"],
           db.data_ptr(), p["rs_b"], p["cs_b"], p["beta"], dc.data_ptr(), p["rs_c"],
           p["cs_c"], ctypes.byref(desc))
    assert st == 0
    torch.cuda.synchronize()
    _check(p, dc.cpu().numpy(), want, dt, case[0])


def test_driver_rejects_bad_descriptors(lib):
    """Enum codes outside the header's values are refused before any work."""
    f = L.driver_fn(lib, np.float64)
    a = np.ones(16)
    c = np.zeros(16)
    for bad in ([3, 0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 4, 0, 0, 0, 0, 0],
                [0, 0, 0, 0, 0, 3, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 7, 0]):
        desc = L.Product(*bad)
        assert f(4, 4, 4, 1.0, a.ctypes.data, 1, 4, a.ctypes.data, 1, 4, 0.0, c.ctypes.data,
                 1, 4, ctypes.byref(desc)) == M.EINVAL
    assert np.all(c == 0)
